/*
 * emx_oracle.c -- serial CPU restatement of the eMix3d per-RK-stage spatial
 * operator.  TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and
 * the cpu_baseline leg of bench.py; never linked into the product.
 *
 * Every function cites the reference lines it restates and keeps the
 * reference's floating-point operation order where that is cheap to do, so
 * that the Fortran golden vectors are met to the last few ulps.
 * Build: gcc -O2 -fPIC -shared (no -ffast-math: SELECTIVO relies on NaN
 * comparisons being false, solver/sgdm_incomp.f90:382-389).
 */
#include "emx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, k) ((long)(i) + (long)nx * ((long)(j) + (long)ny * (long)(k)))

/* Every scheme needs n >= 5: the C4 matrix (a=(0,1/4,1/4,3), c=(3,1/4,1/4,0))
 * is singular at n = 4, where the reference's derivnper divides by zero. */
static int scheme_min_n(int scheme) { (void)scheme; return 5; }

/* Tridiagonal coefficients.
 * C4 : solver/derivada.f90:22-62        a=(0,1/4..,3), c=(3,1/4..,0), b=1
 * C6F: src/MSVSingle/derivada_incomp.f90:20-62
 *      a=(0,1/4,1/3..,1/4,2), c=(2,1/4,1/3..,1/4,0), b=1
 * C6C: src/Schemes/derivCompact.cu:165-189
 *      a=(0,1/4,1/3..,1/4,3), c=(3,1/4,1/3..,1/4,0), b=1 */
int orc_coeffs(int scheme, int n, double *a, double *b, double *c)
{
    if (n < scheme_min_n(scheme)) return -1;
    for (int l = 0; l < n; l++) b[l] = 1.0;
    if (scheme == ORC_C4) {
        for (int l = 1; l < n - 1; l++) { a[l] = 1. / 4.; c[l] = 1. / 4.; }
        a[0] = 0.0; a[n - 1] = 3.0;
        c[0] = 3.0; c[n - 1] = 0.0;
    } else {
        double edge = (scheme == ORC_C6F) ? 2.0 : 3.0;
        for (int l = 2; l < n - 2; l++) { a[l] = 1.0 / 3.0; c[l] = 1.0 / 3.0; }
        a[0] = 0.0; a[1] = 0.25; a[n - 2] = 0.25; a[n - 1] = edge;
        c[0] = edge; c[1] = 0.25; c[n - 2] = 0.25; c[n - 1] = 0.0;
    }
    return 0;
}

/* Right-hand side of one line, written to r[0..n-1].
 * C4 : solver/derivada.f90:104-109 with consrs from :83-89
 * C6F: src/MSVSingle/derivada_incomp.f90:178-185 with consrs from :81-86
 * C6C: src/Schemes/derivCompact.cu:90-140 (the C side multiplies by the
 *      inverse spacing; the Fortran passes 1/deltax and the kernel takes
 *      1.0/delta, so the factor is the same inverse spacing). */
static void rhs_line(int scheme, int n, const double *f, long is, double ih,
                     double *r)
{
#define F(l) f[(long)(l) * is]
    if (scheme == ORC_C4) {
        const double ars = 3. / 4., ars1 = -17. / 6., brs1 = 3. / 2., crs1 = 1.5;
        for (int l = 1; l < n - 1; l++) r[l] = ih * (ars * (F(l + 1) - F(l - 1)));
        r[0] = ih * (ars1 * F(0) + brs1 * F(1) + crs1 * F(2) - 1. / 6. * F(3));
        r[n - 1] = ih * (-ars1 * F(n - 1) - brs1 * F(n - 2) - crs1 * F(n - 3)
                         + 1. / 6. * F(n - 4));
    } else if (scheme == ORC_C6F) {
        const double ars = 7. / 9., brs = 1. / 36., ars1 = -2.5, brs1 = 2.,
                     crs1 = 0.5, ars2 = 0.75;
        for (int l = 2; l < n - 2; l++)
            r[l] = ih * (ars * (F(l + 1) - F(l - 1)) + brs * (F(l + 2) - F(l - 2)));
        r[0] = ih * (ars1 * F(0) + brs1 * F(1) + crs1 * F(2));
        r[n - 1] = ih * (-ars1 * F(n - 1) - brs1 * F(n - 2) - crs1 * F(n - 3));
        r[1] = ih * (ars2 * (F(2) - F(0)));
        r[n - 2] = ih * (ars2 * (F(n - 1) - F(n - 3)));
    } else {
        const double ars = 14.0 / (9.0 * 2.0), brs = 1.0 / (9.0 * 4.0),
                     ars1 = -17.0 / 6.0, brs1 = 3.0 / 2.0, crs1 = 3.0 / 2.0,
                     drs1 = -1.0 / 6.0, ars2 = -3.0 / 4.0, brs2 = 0.0,
                     crs2 = 3.0 / 4.0, drs2 = 0.0;
        for (int l = 2; l < n - 2; l++)
            r[l] = ih * (ars * (F(l + 1) - F(l - 1)) + (brs * (F(l + 2) - F(l - 2))));
        r[0] = ih * (ars1 * F(0) + brs1 * F(1) + crs1 * F(2) + drs1 * F(3));
        r[1] = ih * (ars2 * F(0) + brs2 * F(1) + crs2 * F(2) + drs2 * F(3));
        r[n - 1] = ih * (-ars1 * F(n - 1) - brs1 * F(n - 2) - crs1 * F(n - 3)
                         - drs1 * F(n - 4));
        r[n - 2] = ih * (-ars2 * F(n - 1) - brs2 * F(n - 2) - crs2 * F(n - 3)
                         - drs2 * F(n - 4));
    }
#undef F
}

/* Thomas solve, identical on CPU and GPU in the reference:
 * solver/derivada.f90:111-122, src/Schemes/derivCompact.cu:199-212. */
int orc_deriv_line(int scheme, int n, const double *f, long is, double inv_h,
                   double *out, long os)
{
    if (n < scheme_min_n(scheme)) return -1;
    double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)n);
    double *a = buf, *b = buf + n, *c = buf + 2 * n, *r = buf + 3 * n,
           *gam = buf + 4 * n;
    orc_coeffs(scheme, n, a, b, c);
    rhs_line(scheme, n, f, is, inv_h, r);
    double *u = r; /* solve in place on r, as derivnper does on u */
    double bet = b[0];
    u[0] = r[0] / bet;
    for (int l = 1; l < n; l++) {
        gam[l] = c[l - 1] / bet;
        bet = b[l] - a[l] * gam[l];
        u[l] = (r[l] - a[l] * u[l - 1]) / bet;
    }
    for (int l = n - 2; l >= 0; l--) u[l] = u[l] - gam[l + 1] * u[l + 1];
    for (int l = 0; l < n; l++) out[(long)l * os] = u[l];
    free(buf);
    return 0;
}

int orc_deriv3d(int scheme, int dir, int nx, int ny, int nz, const double *f,
                double inv_h, double *df)
{
    long sx = 1, sy = nx, sz = (long)nx * ny;
    if (dir == 0) {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                if (orc_deriv_line(scheme, nx, f + IDX(0, j, k), sx, inv_h,
                                   df + IDX(0, j, k), sx)) return -1;
    } else if (dir == 1) {
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                if (orc_deriv_line(scheme, ny, f + IDX(i, 0, k), sy, inv_h,
                                   df + IDX(i, 0, k), sy)) return -1;
    } else {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                if (orc_deriv_line(scheme, nz, f + IDX(i, j, 0), sz, inv_h,
                                   df + IDX(i, j, 0), sz)) return -1;
    }
    return 0;
}

/* solver/jacob_malla.f90:14-169.  Component c of jbn is at jbn + c*N
 * (Fortran component c+1).  Derivative order per jacob_malla: x_xi->1,
 * y_xi->4, z_xi->7, x_eta->2, y_eta->5, z_eta->8, x_zeta->3, y_zeta->6,
 * z_zeta->9; then det -> 10, 1/det -> 11, inverse (with its sign pattern)
 * -> 1..9.  Returns -2 where |det| <= 1e-20 (the reference STOPs,
 * jacob_malla.f90:136-139). */
int orc_jacobeano(int scheme, int nx, int ny, int nz, const double *x,
                  double *jbn)
{
    long N = (long)nx * ny * nz;
    double ih[3] = {(double)(nx - 1), (double)(ny - 1), (double)(nz - 1)};
    /* (coordinate, direction) -> Fortran component */
    static const int comp[3][3] = {{1, 2, 3}, {4, 5, 6}, {7, 8, 9}};
    for (int xc = 0; xc < 3; xc++)
        for (int d = 0; d < 3; d++)
            if (orc_deriv3d(scheme, d, nx, ny, nz, x + xc * N, ih[d],
                            jbn + (comp[xc][d] - 1) * N)) return -1;
    for (long p = 0; p < N; p++) {
#define J(c) jbn[((c) - 1) * N + p]
        double det = J(1) * J(5) * J(9) - J(1) * J(6) * J(8) + J(2) * J(6) * J(7)
                   - J(2) * J(4) * J(9) + J(3) * J(4) * J(8) - J(3) * J(5) * J(7);
        J(10) = det;
        if (fabs(det) <= 1.e-20) return -2;
    }
    for (long p = 0; p < N; p++) {
        J(11) = 1. / J(10);
        double t1 = J(5) * J(9) - J(6) * J(8);
        double t4 = J(3) * J(8) - J(2) * J(9);
        double t7 = J(2) * J(6) - J(3) * J(5);
        double t2 = J(6) * J(7) - J(4) * J(9);
        double t5 = J(1) * J(9) - J(3) * J(7);
        double t8 = J(3) * J(4) - J(1) * J(6);
        double t3 = J(4) * J(8) - J(5) * J(7);
        double t6 = J(2) * J(7) - J(1) * J(8);
        double t9 = J(1) * J(5) - J(2) * J(4);
        double r = J(11);
        J(1) = t1 * r;
        J(2) = -t2 * r;
        J(3) = t3 * r;
        J(4) = -t4 * r;
        J(5) = t5 * r;
        J(6) = -t6 * r;
        J(7) = t7 * r;
        J(8) = -t8 * r;
        J(9) = t9 * r;
#undef J
    }
    return 0;
}

/* solver/deriv_vel.f90:17-161: dvel(1..9) = u_xi, v_xi, w_xi, u_eta, v_eta,
 * w_eta, u_zeta, v_zeta, w_zeta; dtemp(1..3) = T_xi, T_eta, T_zeta. */
int orc_deriv_vel(int scheme, int nx, int ny, int nz, const double *u,
                  const double *temp, double *dvel, double *dtemp,
                  double ihx, double ihy, double ihz)
{
    long N = (long)nx * ny * nz;
    double ih[3] = {ihx, ihy, ihz};
    for (int d = 0; d < 3; d++)
        for (int c = 0; c < 3; c++)
            if (orc_deriv3d(scheme, d, nx, ny, nz, u + c * N, ih[d],
                            dvel + (3 * d + c) * N)) return -1;
    if (temp && dtemp)
        for (int d = 0; d < 3; d++)
            if (orc_deriv3d(scheme, d, nx, ny, nz, temp, ih[d], dtemp + d * N))
                return -1;
    return 0;
}

/* solver/flujos_compact.f90:20-217.  Fortran jbn(:,:,:,1)=xi_x -> J1 etc;
 * dvel(:,:,:,m) -> D(m). */
int orc_flujos(int neq, int nx, int ny, int nz, const double *um,
               const double *u, const double *pres, const double *vis,
               const double *jbn, const double *dvel, double *flux)
{
    long N = (long)nx * ny * nz;
    if (neq < 1 || neq > 5) return -1;
    double *e = flux, *f = flux + N, *g = flux + 2 * N, *ev = flux + 3 * N,
           *fv = flux + 4 * N, *gv = flux + 5 * N;
    for (long p = 0; p < N; p++) {
#define UM(m) um[((m) - 1) * N + p]
#define U(m) u[((m) - 1) * N + p]
#define J(c) jbn[((c) - 1) * N + p]
#define D(c) dvel[((c) - 1) * N + p]
        double mu = vis[p];
        switch (neq) {
        case 1:
            e[p] = -UM(2); f[p] = -UM(3); g[p] = -UM(4);
            ev[p] = 0.0; fv[p] = 0.0; gv[p] = 0.0;
            break;
        case 2:
            e[p] = -UM(2) * U(1) - pres[p];
            f[p] = -UM(3) * U(1);
            g[p] = -UM(4) * U(1);
            ev[p] = mu * (2.0 * (J(1) * D(1)));
            fv[p] = mu * (J(1) * D(2) + J(5) * D(4));
            gv[p] = mu * (J(1) * D(3) + J(9) * D(7));
            break;
        case 3:
            e[p] = -UM(2) * U(2);
            f[p] = -UM(3) * U(2) - pres[p];
            g[p] = -UM(4) * U(2);
            ev[p] = mu * (J(1) * D(2) + J(5) * D(4));
            fv[p] = mu * (2.0 * (J(5) * D(5)));
            gv[p] = mu * (J(9) * D(8) + J(5) * D(6));
            break;
        case 4:
            e[p] = -UM(2) * U(3);
            f[p] = -UM(3) * U(3);
            g[p] = -UM(4) * U(3) - pres[p];
            ev[p] = mu * (J(1) * D(3) + J(9) * D(7));
            fv[p] = mu * (J(9) * D(8) + J(5) * D(6));
            gv[p] = mu * (2.0 * (J(9) * D(9)));
            break;
        case 5:
            e[p] = -U(1) * UM(5); f[p] = -U(2) * UM(5); g[p] = -U(3) * UM(5);
            ev[p] = 0.0; fv[p] = 0.0; gv[p] = 0.0;
            break;
        }
#undef UM
#undef U
#undef D
    }
    return 0;
}

/* solver/divergencia.f90:225-354: project E=J*(xi_x*e) etc., differentiate
 * along the matching direction, rs = (1/J)*(E_xi + F_eta + G_zeta). */
int orc_divergencia(int scheme, int nx, int ny, int nz, const double *flux,
                    const double *jbn, double ihx, double ihy, double ihz,
                    double *rs, double *rsv)
{
    long N = (long)nx * ny * nz;
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)N);
    double *d = (double *)malloc(sizeof(double) * 6 * (size_t)N);
    if (!w || !d) { free(w); free(d); return -3; }
    static const int mcomp[3] = {1, 5, 9};
    double ih[3] = {ihx, ihy, ihz};
    int rc = 0;
    for (int q = 0; q < 6 && !rc; q++) {
        int dir = q % 3;
        const double *src = flux + q * N;
        const double *met = jbn + (mcomp[dir] - 1) * N;
        const double *det = jbn + 9 * N;
        for (long p = 0; p < N; p++) w[p] = det[p] * (met[p] * src[p]);
        rc = orc_deriv3d(scheme, dir, nx, ny, nz, w, ih[dir], d + q * N);
    }
    if (!rc) {
        const double *ij = jbn + 10 * N;
        for (long p = 0; p < N; p++) {
            rs[p] = ij[p] * (d[p] + d[N + p] + d[2 * N + p]);
            rsv[p] = ij[p] * (d[3 * N + p] + d[4 * N + p] + d[5 * N + p]);
        }
    }
    free(w);
    free(d);
    return rc;
}

/* solver/sgdm_incomp.f90:20-108 (INISGDM). */
int orc_inisgdm(int nx, int ny, int nz, const double *x, double *sgd,
                double *dmins, double *deltamax)
{
    long N = (long)nx * ny * nz;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
    const double *X = x, *Y = x + N, *Z = x + 2 * N;
    double *dxyz = sgd, *dxp = sgd + N, *dxm = sgd + 2 * N, *dyp = sgd + 3 * N,
           *dym = sgd + 4 * N, *dzp = sgd + 5 * N, *dzm = sgd + 6 * N;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long p = IDX(i, j, k);
                int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1,
                    km = k - 1;
                if (i == 0) { dxp[p] = X[IDX(1, j, k)] - X[IDX(0, j, k)]; dxm[p] = dxp[p]; }
                else if (i == nx - 1) { dxm[p] = X[IDX(nx - 1, j, k)] - X[IDX(nx - 2, j, k)]; dxp[p] = dxm[p]; }
                else { dxm[p] = X[p] - X[IDX(im, j, k)]; dxp[p] = X[IDX(ip, j, k)] - X[p]; }
                if (j == 0) { dyp[p] = Y[IDX(i, 1, k)] - Y[IDX(i, 0, k)]; dym[p] = dyp[p]; }
                else if (j == ny - 1) { dym[p] = Y[IDX(i, ny - 1, k)] - Y[IDX(i, ny - 2, k)]; dyp[p] = dym[p]; }
                else { dym[p] = Y[p] - Y[IDX(i, jm, k)]; dyp[p] = Y[IDX(i, jp, k)] - Y[p]; }
                if (k == 0) { dzp[p] = Z[IDX(i, j, 1)] - Z[IDX(i, j, 0)]; dzm[p] = dzp[p]; }
                else if (k == nz - 1) { dzm[p] = Z[IDX(i, j, nz - 1)] - Z[IDX(i, j, nz - 2)]; dzp[p] = dzm[p]; }
                else { dzm[p] = Z[p] - Z[IDX(i, j, km)]; dzp[p] = Z[IDX(i, j, kp)] - Z[p]; }
            }
    double dmax = 0.0;
    for (long p = 0; p < N; p++) {
        double dzmax = fmax(dzm[p], dzp[p]);
        double dymax = fmax(dym[p], dyp[p]);
        double dxmax = fmax(dxm[p], dxp[p]);
        double dzmin = fmin(dzm[p], dzp[p]);
        double dymin = fmin(dym[p], dyp[p]);
        double dxmin = fmin(dxm[p], dxp[p]);
        double t = fmax(fmax(dxmax, dymax), fmax(dzmax, dmax));
        dmax = t;
        if (dmins) dmins[p] = fmin(fmin(dzmin, dymin), dxmin);
        dxyz[p] = pow(dxmax * dymax * dzmax, 1. / 3.);
    }
    for (long p = 0; p < N; p++) {
        dxp[p] = pow(dxyz[p] / dxp[p], 1.0 / 3.0);
        dxm[p] = pow(dxyz[p] / dxm[p], 1.0 / 3.0);
        dyp[p] = pow(dxyz[p] / dyp[p], 1.0 / 3.0);
        dym[p] = pow(dxyz[p] / dym[p], 1.0 / 3.0);
        dzp[p] = pow(dxyz[p] / dzp[p], 1.0 / 3.0);
        dzm[p] = pow(dxyz[p] / dzm[p], 1.0 / 3.0);
    }
    if (deltamax) *deltamax = dmax;
    return 0;
}

/* viscosidad_t (solver/sgdm_incomp.f90:117-162) = f_structura (:166-315),
 * amut = amu0*sqrt(F2)*dxyzsgd, VORT (:319-343), SELECTIVO (:346-395);
 * then viscosidad (solver/viscosidad_incomp.f90:12-20).
 * isgm=2, amu0=0.1781, cutoff=cos(20 deg) (sgdm_incomp.f90:13-19). */
int orc_sgdm(int nx, int ny, int nz, const double *u, const double *dvel,
             const double *jbn, const double *sgd, double c4, double *amut,
             double *vis, double *vis5)
{
    long N = (long)nx * ny * nz;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
    const double amu0 = 0.1781;
    const double cutoff = cos(20.0 * atan(1.0) / 45.0);
    const double *dxyz = sgd, *dxp = sgd + N, *dxm = sgd + 2 * N,
                 *dyp = sgd + 3 * N, *dym = sgd + 4 * N, *dzp = sgd + 5 * N,
                 *dzm = sgd + 6 * N;
    const double *U1 = u, *U2 = u + N, *U3 = u + 2 * N;
    /* f_structura: six one-sided squared differences, accumulated in the
     * reference's loop order x+, x-, y+, y-, z+, z-. */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long p = IDX(i, j, k);
                long a, b;
                double s, du, dv, dw;
                /* x+ (i<nx-1: (i+1)-(i); i=nx-1: (nx-1)-(nx-2)) */
                if (i < nx - 1) { a = IDX(i + 1, j, k); b = p; }
                else { a = IDX(nx - 1, j, k); b = IDX(nx - 2, j, k); }
                du = (U1[a] - U1[b]) * dxp[p]; dv = (U2[a] - U2[b]) * dxp[p];
                dw = (U3[a] - U3[b]) * dxp[p];
                s = du * du + dv * dv + dw * dw;
                /* x- (i>0: (i-1)-(i); i=0: (1)-(0)) */
                if (i > 0) { a = IDX(i - 1, j, k); b = p; }
                else { a = IDX(1, j, k); b = IDX(0, j, k); }
                du = (U1[a] - U1[b]) * dxm[p]; dv = (U2[a] - U2[b]) * dxm[p];
                dw = (U3[a] - U3[b]) * dxm[p];
                s = s + du * du + dv * dv + dw * dw;
                /* y+ */
                if (j < ny - 1) { a = IDX(i, j + 1, k); b = p; }
                else { a = IDX(i, ny - 1, k); b = IDX(i, ny - 2, k); }
                du = (U1[a] - U1[b]) * dyp[p]; dv = (U2[a] - U2[b]) * dyp[p];
                dw = (U3[a] - U3[b]) * dyp[p];
                s = s + du * du + dv * dv + dw * dw;
                /* y- (j=0: (0)-(1)) */
                if (j > 0) { a = IDX(i, j - 1, k); b = p; }
                else { a = IDX(i, 0, k); b = IDX(i, 1, k); }
                du = (U1[a] - U1[b]) * dym[p]; dv = (U2[a] - U2[b]) * dym[p];
                dw = (U3[a] - U3[b]) * dym[p];
                s = s + du * du + dv * dv + dw * dw;
                /* z+ */
                if (k < nz - 1) { a = IDX(i, j, k + 1); b = p; }
                else { a = IDX(i, j, nz - 1); b = IDX(i, j, nz - 2); }
                du = (U1[a] - U1[b]) * dzp[p]; dv = (U2[a] - U2[b]) * dzp[p];
                dw = (U3[a] - U3[b]) * dzp[p];
                s = s + du * du + dv * dv + dw * dw;
                /* z- (k=0: (0)-(1)) */
                if (k > 0) { a = IDX(i, j, k - 1); b = p; }
                else { a = IDX(i, j, 0); b = IDX(i, j, 1); }
                du = (U1[a] - U1[b]) * dzm[p]; dv = (U2[a] - U2[b]) * dzm[p];
                dw = (U3[a] - U3[b]) * dzm[p];
                s = s + du * du + dv * dv + dw * dw;
                amut[p] = s / 6.e0;
            }
    for (long p = 0; p < N; p++) amut[p] = amu0 * sqrt(amut[p]) * dxyz[p];
    /* VORT, with the reference's metric/derivative pairing (sgdm_incomp.f90:
     * 333-338): wx = jbn5*dvel8 - jbn9*dvel6, wy = jbn9*dvel3 - jbn1*dvel7,
     * wz = jbn1*dvel4 - jbn5*dvel2. */
    double *w = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    if (!w) return -3;
    double *wx = w, *wy = w + N, *wz = w + 2 * N;
    const double *J1 = jbn, *J5 = jbn + 4 * N, *J9 = jbn + 8 * N;
#define D(c) dvel[((c) - 1) * N + p]
    for (long p = 0; p < N; p++) {
        wx[p] = J5[p] * D(8) - J9[p] * D(6);
        wy[p] = J9[p] * D(3) - J1[p] * D(7);
        wz[p] = J1[p] * D(4) - J5[p] * D(2);
    }
#undef D
    /* SELECTIVO: neighbour sum in the order ip, im, kp, km, jp, jm with
     * clamped indices; amut=0 where cos(angle) > cutoff (NaN -> keep). */
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long p = IDX(i, j, k);
                int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1,
                    km = k - 1;
                if (i == nx - 1) ip = nx - 1;
                if (i == 0) im = 0;
                if (j == ny - 1) jp = ny - 1;
                if (j == 0) jm = 0;
                if (k == nz - 1) kp = nz - 1;
                if (k == 0) km = 0;
                long n1 = IDX(ip, j, k), n2 = IDX(im, j, k), n3 = IDX(i, j, kp),
                     n4 = IDX(i, j, km), n5 = IDX(i, jp, k), n6 = IDX(i, jm, k);
                double wxn = wx[n1] + wx[n2] + wx[n3] + wx[n4] + wx[n5] + wx[n6];
                double wyn = wy[n1] + wy[n2] + wy[n3] + wy[n4] + wy[n5] + wy[n6];
                double wzn = wz[n1] + wz[n2] + wz[n3] + wz[n4] + wz[n5] + wz[n6];
                double wn = wx[p] * wx[p] + wy[p] * wy[p] + wz[p] * wz[p];
                wn = sqrt(wn * (wxn * wxn + wyn * wyn + wzn * wzn));
                double ps1 = wx[p] * wxn, ps2 = wy[p] * wyn, ps3 = wz[p] * wzn;
                double s1 = (ps1 + ps2 + ps3) / wn;
                if (s1 > cutoff) amut[p] = 0.0;
            }
    free(w);
    for (long p = 0; p < N; p++) {
        if (vis5) vis5[p] = amut[p] / 0.7;
        vis[p] = c4 * 1.0 + amut[p];
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Level set.                                                              */

void orc_pad_mirror(int nx, int ny, int nz, int g, const double *in,
                    double *out)
{
    int ngx = nx + 2 * g, ngy = ny + 2 * g, ngz = nz + 2 * g;
    for (int k = 0; k < ngz; k++) {
        int ks = k < g ? 2 * g - 1 - k : (k >= nz + g ? 2 * (nz + g) - 1 - k : k);
        ks -= g;
        for (int j = 0; j < ngy; j++) {
            int js = j < g ? 2 * g - 1 - j : (j >= ny + g ? 2 * (ny + g) - 1 - j : j);
            js -= g;
            for (int i = 0; i < ngx; i++) {
                int is = i < g ? 2 * g - 1 - i : (i >= nx + g ? 2 * (nx + g) - 1 - i : i);
                is -= g;
                out[(long)i + (long)ngx * ((long)j + (long)ngy * k)] =
                    in[IDX(is, js, ks)];
            }
        }
    }
}

/* DevFirstOrder_LS, src/LevelSet/extrapol.cu:105-153 (factor 1 on any
 * boundary point, 0.5 elsewhere; one-sided on the boundary). */
static void ls_grad_first_order(int Nx, int Ny, int Nz, const double *phiS,
                                const double *jbn, double dX, double dY,
                                double dZ, double *dphi)
{
    long Off = (long)Nx * Ny * Nz;
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                long id = (long)Nx * Ny * z + (long)Nx * y + x;
                long ip = id + 1, im = id - 1, jp = id + Nx, jm = id - Nx,
                     kp = id + (long)Nx * Ny, km = id - (long)Nx * Ny;
                double factor = 0.5;
                if (x == 0) { im = id; factor = 1.0; }
                if (y == 0) { jm = id; factor = 1.0; }
                if (z == 0) { km = id; factor = 1.0; }
                if (x == Nx - 1) { ip = id; factor = 1.0; }
                if (y == Ny - 1) { jp = id; factor = 1.0; }
                if (z == Nz - 1) { kp = id; factor = 1.0; }
                dphi[id] = factor * dX * jbn[id] * (phiS[ip] - phiS[im]);
                dphi[Off + id] = factor * dY * jbn[id + 4 * Off] * (phiS[jp] - phiS[jm]);
                dphi[2 * Off + id] = factor * dZ * jbn[id + 8 * Off] * (phiS[kp] - phiS[km]);
            }
}

/* extrapolKernel, src/LevelSet/extrapol.cu:5-103, with extFlag > 0. */
static void ls_extrap_rhs(int Nx, int Ny, int Nz, double *rs,
                          const double *ext, const double *phiS,
                          const double *jbn, const double *dphi, double dX,
                          double dY, double dZ)
{
    long Off = (long)Nx * Ny * Nz;
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                long id = (long)Nx * Ny * z + (long)Nx * y + x;
                long ip = id + 1, im = id - 1, jp = id + Nx, jm = id - Nx,
                     kp = id + (long)Nx * Ny, km = id - (long)Nx * Ny;
                long id2 = id;
                if (x == 0) { id2 = ip; im = id; }
                if (y == 0) { id2 = jp; jm = id; }
                if (z == 0) { id2 = kp; km = id; }
                if (x == Nx - 1) { id2 = im; ip = id; }
                if (y == Ny - 1) { id2 = jm; jp = id; }
                if (z == Nz - 1) { id2 = km; kp = id; }
                double so = (double)(phiS[id] > 0.0);
                double pdx = so * dphi[id], pdy = so * dphi[id + Off],
                       pdz = so * dphi[id + 2 * Off];
                double xd = dX * jbn[id] * (ext[id2] - ext[im]);
                double yd = dY * jbn[id + 4 * Off] * (ext[id2] - ext[jm]);
                double zd = dZ * jbn[id + 8 * Off] * (ext[id2] - ext[km]);
                double xu = dX * jbn[id] * (ext[ip] - ext[id2]);
                double yu = dY * jbn[id + 4 * Off] * (ext[jp] - ext[id2]);
                double zu = dZ * jbn[id + 8 * Off] * (ext[kp] - ext[id2]);
                double s = (double)(pdx > 0.0) - (double)(pdx < 0.0);
                double r = (0.5 * (s + 1.0) * xd + 0.5 * fabs(s - 1.0) * xu) * pdx;
                s = (double)(pdy > 0.0) - (double)(pdy < 0.0);
                r += (0.5 * (s + 1.0) * yd + 0.5 * fabs(s - 1.0) * yu) * pdy;
                s = (double)(pdz > 0.0) - (double)(pdz < 0.0);
                r += (0.5 * (s + 1.0) * zd + 0.5 * fabs(s - 1.0) * zu) * pdz;
                rs[id] = r;
            }
}

int orc_ls_extrapolate(int nx, int ny, int nz, double *pres, const double *jbn,
                       const double *sobject, double dx, double dy, double dz,
                       double deltamin)
{
    const int g = 5; /* gcells, lsTransportEc.cu:39 */
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    int gx = nx + 2 * g, gy = ny + 2 * g, gz = nz + 2 * g;
    long N = (long)nx * ny * nz, NG = (long)gx * gy * gz;
    double *P = (double *)malloc(sizeof(double) * (size_t)NG);
    double *S = (double *)malloc(sizeof(double) * (size_t)NG);
    double *JB = (double *)calloc((size_t)(11 * NG), sizeof(double));
    double *dphi = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *rs = (double *)malloc(sizeof(double) * (size_t)NG);
    double *d = (double *)malloc(sizeof(double) * (size_t)NG);
    if (!P || !S || !JB || !dphi || !rs || !d) {
        free(P); free(S); free(JB); free(dphi); free(rs); free(d);
        return -3;
    }
    orc_pad_mirror(nx, ny, nz, g, pres, P);
    orc_pad_mirror(nx, ny, nz, g, sobject, S);
    orc_pad_mirror(nx, ny, nz, g, jbn, JB);
    orc_pad_mirror(nx, ny, nz, g, jbn + 4 * N, JB + 4 * NG);
    orc_pad_mirror(nx, ny, nz, g, jbn + 8 * N, JB + 8 * NG);
    /* cuExtrapolation(..., deltaX, ..., deltamin*0.5, 1): kernels receive
     * 1.0/deltaX (interface.cu:373-421). */
    double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz, dt = deltamin * 0.5;
    ls_grad_first_order(gx, gy, gz, S, JB, iX, iY, iZ, dphi);
    for (int it = 1; it <= 10; it++) {
        ls_extrap_rhs(gx, gy, gz, rs, P, S, JB, dphi, iX, iY, iZ);
        for (long p = 0; p < NG; p++) d[p] = P[p] - dt * rs[p];
        ls_extrap_rhs(gx, gy, gz, rs, d, S, JB, dphi, iX, iY, iZ);
        for (long p = 0; p < NG; p++) d[p] = 0.75 * P[p] + 0.25 * (d[p] - dt * rs[p]);
        ls_extrap_rhs(gx, gy, gz, rs, d, S, JB, dphi, iX, iY, iZ);
        for (long p = 0; p < NG; p++) P[p] = (P[p] + 2.0 * (d[p] - dt * rs[p])) / 3.0;
    }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                pres[IDX(i, j, k)] =
                    P[(long)(i + g) + (long)gx * ((long)(j + g) + (long)gy * (k + g))];
    free(P); free(S); free(JB); free(dphi); free(rs); free(d);
    return 0;
}

/* Phi_x_WENO, src/LevelSet/lsTools.cu:9-50 (WENO5-JS, eps=1e-6). */
static double weno5(double b1, double b2, double b3, double b4, double b5)
{
    double s1 = (13.0 / 12.0) * (b1 - 2.0 * b2 + b3) * (b1 - 2.0 * b2 + b3)
              + (0.25) * (b1 - 4.0 * b2 + 3.0 * b3) * (b1 - 4.0 * b2 + 3.0 * b3);
    double s2 = (13.0 / 12.0) * (b2 - 2.0 * b3 + b4) * (b2 - 2.0 * b3 + b4)
              + (0.25) * (b2 - b4) * (b2 - b4);
    double s3 = (13.0 / 12.0) * (b3 - 2.0 * b4 + b5) * (b3 - 2.0 * b4 + b5)
              + (0.25) * (3.0 * b3 - 4.0 * b4 + b5) * (3.0 * b3 - 4.0 * b4 + b5);
    double a1 = 0.1 / ((s1 + 1.0e-6) * (s1 + 1.0e-6));
    double a2 = 0.6 / ((s2 + 1.0e-6) * (s2 + 1.0e-6));
    double a3 = 0.3 / ((s3 + 1.0e-6) * (s3 + 1.0e-6));
    double o1 = a1 / (a1 + a2 + a3);
    double o2 = a2 / (a1 + a2 + a3);
    double o3 = a3 / (a1 + a2 + a3);
    return ((o1 * (2.0 * b1 - 7.0 * b2 + 11.0 * b3)
             + o2 * (-1.0 * b2 + 5.0 * b3 + 2.0 * b4)
             + o3 * (2.0 * b3 + 5.0 * b4 - b5)) * (1.0 / 6.0));
}

/* Dev1thO_Downwind, lsTools.cu:53-95: backward differences, index 0 uses
 * phi(1)-phi(0). */
static void ls_backward_diff(int Nx, int Ny, int Nz, double *dphi,
                             const double *phi, double dX, double dY,
                             double dZ)
{
    long Off = (long)Nx * Ny * Nz;
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                long id = (long)Nx * Ny * z + (long)Nx * y + x;
                long im = id - 1, jm = id - Nx, km = id - (long)Nx * Ny;
                long ix = id, iy = id, iz = id;
                if (x == 0) { im = id; ix = (long)Nx * Ny * z + (long)Nx * y + 1; }
                if (y == 0) { jm = id; iy = (long)Nx * Ny * z + (long)Nx * 1 + x; }
                if (z == 0) { km = id; iz = (long)Nx * Ny * 1 + (long)Nx * y + x; }
                dphi[id] = dX * (phi[ix] - phi[im]);
                dphi[Off + id] = dY * (phi[iy] - phi[jm]);
                dphi[2 * Off + id] = dZ * (phi[iz] - phi[km]);
            }
}

/* PhiDevPlusParameterJB (lsTools.cu:587-673): the left-biased WENO
 * derivative from backward differences at i-2..i+2, scaled by the metric at
 * the centre.  Boundary substitutions as written, including
 * id_kp2 = id_jm1 at the last z plane (lsTools.cu:636). */
static void ls_weno_left(int Nx, int Ny, int Nz, double *out,
                         const double *dphi, const double *jbn)
{
    long Off = (long)Nx * Ny * Nz, sxy = (long)Nx * Ny;
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                long id = sxy * z + (long)Nx * y + x;
                long im1 = id - 1, ip1 = id + 1, jm1 = id - Nx, jp1 = id + Nx,
                     km1 = id - sxy, kp1 = id + sxy, im2 = id - 2, ip2 = id + 2,
                     jm2 = id - 2 * Nx, jp2 = id + 2 * Nx, km2 = id - 2 * sxy,
                     kp2 = id + 2 * sxy;
                if (x == 0) { im1 = id; im2 = ip1; }
                if (x == 1) { im2 = im1; }
                if (x == Nx - 1) { ip1 = id; ip2 = im1; }
                if (x == Nx - 2) { ip2 = ip1; }
                if (y == 0) { jm1 = id; jm2 = jp1; }
                if (y == 1) { jm2 = jm1; }
                if (y == Ny - 1) { jp1 = id; jp2 = jm1; }
                if (y == Ny - 2) { jp2 = jp1; }
                if (z == 0) { km1 = id; km2 = kp1; }
                if (z == 1) { km2 = km1; }
                if (z == Nz - 1) { kp1 = id; kp2 = jm1; }
                if (z == Nz - 2) { kp2 = kp1; }
                double m = jbn[id];
                out[id] = weno5(dphi[im2] * m, dphi[im1] * m, dphi[id] * m,
                                dphi[ip1] * m, dphi[ip2] * m);
                const double *d1 = dphi + Off;
                m = jbn[id + 4 * Off];
                out[id + Off] = weno5(d1[jm2] * m, d1[jm1] * m, d1[id] * m,
                                      d1[jp1] * m, d1[jp2] * m);
                const double *d2 = dphi + 2 * Off;
                m = jbn[id + 8 * Off];
                out[id + 2 * Off] = weno5(d2[km2] * m, d2[km1] * m, d2[id] * m,
                                          d2[kp1] * m, d2[kp2] * m);
            }
}

/* PhiDevMinusParameterJB (lsTools.cu:676-770): the right-biased WENO
 * derivative from backward differences at i+3..i-1. */
static void ls_weno_right(int Nx, int Ny, int Nz, double *out,
                          const double *dphi, const double *jbn)
{
    long Off = (long)Nx * Ny * Nz, sxy = (long)Nx * Ny;
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int x = 0; x < Nx; x++) {
                long id = sxy * z + (long)Nx * y + x;
                long im1 = id - 1, im2 = id - 2, ip1 = id + 1, jm1 = id - Nx,
                     jm2 = id - 2 * Nx, jp1 = id + Nx, km1 = id - sxy,
                     km2 = id - 2 * sxy, kp1 = id + sxy, ip2 = id + 2,
                     jp2 = id + 2 * Nx, kp2 = id + 2 * sxy, ip3 = id + 3,
                     jp3 = id + 3 * Nx, kp3 = id + 3 * sxy;
                if (x == 0) { im1 = id; }
                if (x == Nx - 1) { ip1 = id; ip2 = im1; ip3 = im2; }
                if (x == Nx - 2) { ip2 = ip1; ip3 = id; }
                if (x == Nx - 3) { ip3 = ip2; }
                if (y == 0) { jm1 = id; }
                if (y == Ny - 1) { jp1 = id; jp2 = jm1; jp3 = jm2; }
                if (y == Ny - 2) { jp2 = jp1; jp3 = id; }
                if (y == Ny - 3) { jp3 = jp2; }
                if (z == 0) { km1 = id; }
                if (z == Nz - 1) { kp1 = id; kp2 = km1; kp3 = km2; }
                if (z == Nz - 2) { kp2 = kp1; kp3 = id; }
                if (z == Nz - 3) { kp3 = kp2; }
                double m = jbn[id];
                out[id] = weno5(dphi[ip3] * m, dphi[ip2] * m, dphi[ip1] * m,
                                dphi[id] * m, dphi[im1] * m);
                const double *d1 = dphi + Off;
                m = jbn[id + 4 * Off];
                out[id + Off] = weno5(d1[jp3] * m, d1[jp2] * m, d1[jp1] * m,
                                      d1[id] * m, d1[jm1] * m);
                const double *d2 = dphi + 2 * Off;
                m = jbn[id + 8 * Off];
                out[id + 2 * Off] = weno5(d2[kp3] * m, d2[kp2] * m, d2[kp1] * m,
                                          d2[id] * m, d2[km1] * m);
            }
}

/* advect_RS_WENO, lsTools.cu:361-417: upwind selection by the sign of each
 * velocity component. */
static void ls_advect_rhs(int Nx, int Ny, int Nz, double *rs,
                          const double *vel, const double *pl,
                          const double *pr)
{
    long Off = (long)Nx * Ny * Nz;
    for (long id = 0; id < Off; id++) {
        double r[3];
        for (int c = 0; c < 3; c++) {
            double v = vel[id + c * Off];
            double s = (double)(v > 0.0) - (double)(v < 0.0);
            r[c] = 0.5 * (s + 1.0) * v * pl[id + c * Off]
                 + 0.5 * fabs(s - 1.0) * v * pr[id + c * Off];
        }
        rs[id] = r[0] + r[1] + r[2];
    }
}

/* The two one-sided WENO5 derivatives of phi on a padded grid (the
 * PhiDevPlus/MinusParameterJB pair after Dev1thO_Downwind), 3 comps each. */
int orc_weno5_derivs(int Nx, int Ny, int Nz, const double *phi,
                     const double *jbn, double dx, double dy, double dz,
                     double *left, double *right)
{
    if (Nx < 4 || Ny < 4 || Nz < 4) return -1;
    long NG = (long)Nx * Ny * Nz;
    double *dphi = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    if (!dphi) return -3;
    ls_backward_diff(Nx, Ny, Nz, dphi, phi, 1.0 / dx, 1.0 / dy, 1.0 / dz);
    ls_weno_left(Nx, Ny, Nz, left, dphi, jbn);
    ls_weno_right(Nx, Ny, Nz, right, dphi, jbn);
    free(dphi);
    return 0;
}

int orc_ls_advect_weno5_padded(int Nx, int Ny, int Nz, double *phi,
                               const double *vel, const double *jbn,
                               double dx, double dy, double dz, double dt)
{
    if (Nx < 4 || Ny < 4 || Nz < 4) return -1;
    long NG = (long)Nx * Ny * Nz;
    double *dphi = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *pl = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *pr = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *rs = (double *)malloc(sizeof(double) * (size_t)NG);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)NG);
    if (!dphi || !pl || !pr || !rs || !tmp) {
        free(dphi); free(pl); free(pr); free(rs); free(tmp);
        return -3;
    }
    double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz;
    /* stage 1 (lsTransportEc.cu:244-292) */
    ls_backward_diff(Nx, Ny, Nz, dphi, phi, iX, iY, iZ);
    ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
    ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
    ls_advect_rhs(Nx, Ny, Nz, rs, vel, pl, pr);
    for (long p = 0; p < NG; p++) tmp[p] = phi[p] - dt * rs[p];
    /* stage 2 (:295-346) */
    ls_backward_diff(Nx, Ny, Nz, dphi, tmp, iX, iY, iZ);
    ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
    ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
    ls_advect_rhs(Nx, Ny, Nz, rs, vel, pl, pr);
    for (long p = 0; p < NG; p++) tmp[p] = 0.75 * phi[p] + 0.25 * (tmp[p] - dt * rs[p]);
    /* stage 3 (:349-399) */
    ls_backward_diff(Nx, Ny, Nz, dphi, tmp, iX, iY, iZ);
    ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
    ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
    ls_advect_rhs(Nx, Ny, Nz, rs, vel, pl, pr);
    for (long p = 0; p < NG; p++) phi[p] = (phi[p] + 2.0 * (tmp[p] - dt * rs[p])) / 3.0;
    free(dphi); free(pl); free(pr); free(rs); free(tmp);
    return 0;
}

int orc_ls_advect_weno5(int nx, int ny, int nz, double *phi, const double *vel,
                        const double *jbn, double dx, double dy, double dz,
                        double dt)
{
    const int g = 5;
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    int gx = nx + 2 * g, gy = ny + 2 * g, gz = nz + 2 * g;
    long N = (long)nx * ny * nz, NG = (long)gx * gy * gz;
    double *P = (double *)malloc(sizeof(double) * (size_t)NG);
    double *V = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *JB = (double *)calloc((size_t)(11 * NG), sizeof(double));
    if (!P || !V || !JB) { free(P); free(V); free(JB); return -3; }
    orc_pad_mirror(nx, ny, nz, g, phi, P);
    for (int c = 0; c < 3; c++) orc_pad_mirror(nx, ny, nz, g, vel + c * N, V + c * NG);
    orc_pad_mirror(nx, ny, nz, g, jbn, JB);
    orc_pad_mirror(nx, ny, nz, g, jbn + 4 * N, JB + 4 * NG);
    orc_pad_mirror(nx, ny, nz, g, jbn + 8 * N, JB + 8 * NG);
    int rc = orc_ls_advect_weno5_padded(gx, gy, gz, P, V, JB, dx, dy, dz, dt);
    if (!rc)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    phi[IDX(i, j, k)] =
                        P[(long)(i + g) + (long)gx * ((long)(j + g) + (long)gy * (k + g))];
    free(P); free(V); free(JB);
    return rc;
}

/* cuGhostCellsMirror3dZ, then Y, then X (lsTools.cu:482-585) on a padded
 * array of Nx x Ny x Nz points with g ghost layers, over the full padded
 * ranges: ghost layer g-1-i takes interior layer g+i, layer n+g+i takes
 * n-1-i+g (n = interior points), direction factor 1. */
static void ls_ghost_mirror(int Nx, int Ny, int Nz, int g, double *a)
{
    const int nx = Nx - 2 * g, ny = Ny - 2 * g, nz = Nz - 2 * g;
    const long sxy = (long)Nx * Ny;
    for (int y = 0; y < Ny; y++)
        for (int x = 0; x < Nx; x++)
            for (int i = 0; i < g; i++) {
                a[x + (long)Nx * y + sxy * (g - i - 1)] = a[x + (long)Nx * y + sxy * (g + i)];
                a[x + (long)Nx * y + sxy * (nz + i + g)] = a[x + (long)Nx * y + sxy * (nz - i - 1 + g)];
            }
    for (int z = 0; z < Nz; z++)
        for (int x = 0; x < Nx; x++)
            for (int i = 0; i < g; i++) {
                a[x + (long)Nx * (g - i - 1) + sxy * z] = a[x + (long)Nx * (g + i) + sxy * z];
                a[x + (long)Nx * (ny + i + g) + sxy * z] = a[x + (long)Nx * (ny - i - 1 + g) + sxy * z];
            }
    for (int z = 0; z < Nz; z++)
        for (int y = 0; y < Ny; y++)
            for (int i = 0; i < g; i++) {
                a[(g - i - 1) + (long)Nx * y + sxy * z] = a[(g + i) + (long)Nx * y + sxy * z];
                a[(nx + i + g) + (long)Nx * y + sxy * z] = a[(nx - i - 1 + g) + (long)Nx * y + sxy * z];
            }
}

int orc_ls_ghost_mirror(int Nx, int Ny, int Nz, int g, double *a)
{
    if (Nx <= 2 * g || Ny <= 2 * g || Nz <= 2 * g) return -1;
    ls_ghost_mirror(Nx, Ny, Nz, g, a);
    return 0;
}

/* reini_RS_WENOJB (lsTools.cu:824-900): Godunov |grad phi| from the left
 * (pl, PhiDevPlusParameterJB) and right (pr, PhiDevMinusParameterJB)
 * derivatives by the sign of phi0, rs = S(phi0)*(|grad phi| - 1) with
 * S = phi0/sqrt(phi0^2 + dmin^2). */
static void ls_reinit_rhs(long NG, double *rs, const double *phi0, const double *dmin,
                          const double *pl, const double *pr)
{
    for (long id = 0; id < NG; id++) {
        const double ta = (double)(phi0[id] > 0.0) - (double)(phi0[id] < 0.0);
        double r[3];
        for (int c = 0; c < 3; c++) {
            const double P = pl[id + c * NG], M = pr[id + c * NG];
            const double pp = P > 0.0 ? P : 0.0, mm = M < 0.0 ? M : 0.0;
            double phiMax = pp * pp, phiMin = mm * mm;
            r[c] = 0.5 * (ta + 1.0) * (phiMax > phiMin ? phiMax : phiMin);
            const double mp = M > 0.0 ? M : 0.0, pm = P < 0.0 ? P : 0.0;
            phiMax = mp * mp;
            phiMin = pm * pm;
            r[c] += 0.5 * fabs(ta - 1.0) * (phiMax > phiMin ? phiMax : phiMin);
        }
        const double grad_mod = sqrt(r[0] + r[1] + r[2]);
        const double so = phi0[id] / sqrt(phi0[id] * phi0[id] + dmin[id] * dmin[id]);
        rs[id] = 1.0 * so * (grad_mod - 1.0);
    }
}

/* cuReinitLsJB (src/LevelSet/lsTransportEc.cu:422-676) on a padded grid
 * with g ghost layers: ten iterations of ghost mirroring, phi0 = phi and
 * three SSP-RK3 stages (RunGK_*, extrapol.cu:155-212) with dt = deltamin/2. */
int orc_ls_reinit_padded(int Nx, int Ny, int Nz, int g, double *phi, const double *jbn,
                         const double *dmin, double dx, double dy, double dz, double deltamin)
{
    if (Nx <= 2 * g || Ny <= 2 * g || Nz <= 2 * g || Nx < 4 || Ny < 4 || Nz < 4) return -1;
    long NG = (long)Nx * Ny * Nz;
    double *dphi = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *pl = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *pr = (double *)malloc(sizeof(double) * 3 * (size_t)NG);
    double *rs = (double *)malloc(sizeof(double) * (size_t)NG);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)NG);
    double *phi0 = (double *)malloc(sizeof(double) * (size_t)NG);
    if (!dphi || !pl || !pr || !rs || !tmp || !phi0) {
        free(dphi); free(pl); free(pr); free(rs); free(tmp); free(phi0);
        return -3;
    }
    const double iX = 1.0 / dx, iY = 1.0 / dy, iZ = 1.0 / dz, dt = 0.5 * deltamin;
    for (int it = 1; it <= 10; it++) {
        ls_ghost_mirror(Nx, Ny, Nz, g, phi);
        memcpy(phi0, phi, sizeof(double) * (size_t)NG);
        ls_backward_diff(Nx, Ny, Nz, dphi, phi, iX, iY, iZ);
        ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
        ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
        ls_reinit_rhs(NG, rs, phi0, dmin, pl, pr);
        for (long p = 0; p < NG; p++) tmp[p] = phi[p] - dt * rs[p];
        ls_backward_diff(Nx, Ny, Nz, dphi, tmp, iX, iY, iZ);
        ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
        ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
        ls_reinit_rhs(NG, rs, phi0, dmin, pl, pr);
        for (long p = 0; p < NG; p++) tmp[p] = 0.75 * phi[p] + 0.25 * (tmp[p] - dt * rs[p]);
        ls_backward_diff(Nx, Ny, Nz, dphi, tmp, iX, iY, iZ);
        ls_weno_right(Nx, Ny, Nz, pr, dphi, jbn);
        ls_weno_left(Nx, Ny, Nz, pl, dphi, jbn);
        ls_reinit_rhs(NG, rs, phi0, dmin, pl, pr);
        for (long p = 0; p < NG; p++) phi[p] = (phi[p] + 2.0 * (tmp[p] - dt * rs[p])) / 3.0;
    }
    free(dphi); free(pl); free(pr); free(rs); free(tmp); free(phi0);
    return 0;
}

/* heviside, solver/advectLS.f90:29-59: PI = 2 acos(0) and eps = 1.5
 * deltamax (:36-37), the three cases in the reference's order (:42-51). */
int orc_heviside(long n, const double *phi, double deltamax, double band, double *hevi)
{
    const double pi = 2.0 * acos(0.0), eps = band * deltamax;
    if (n < 0) return -1;
    for (long p = 0; p < n; p++) {
        const double f = phi[p];
        if (f < -eps) hevi[p] = 0.0;
        if (f >= -eps && f <= eps) hevi[p] = 0.5 + f / (2.0 * eps) + sin(pi * f / eps) / (2.0 * pi);
        if (f > eps) hevi[p] = 1.0;
    }
    return 0;
}

int orc_ls_reinit(int nx, int ny, int nz, double *phi, const double *jbn, const double *dmin,
                  double dx, double dy, double dz, double deltamin)
{
    const int g = 5;
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    int gx = nx + 2 * g, gy = ny + 2 * g, gz = nz + 2 * g;
    long N = (long)nx * ny * nz, NG = (long)gx * gy * gz;
    double *P = (double *)malloc(sizeof(double) * (size_t)NG);
    double *D = (double *)malloc(sizeof(double) * (size_t)NG);
    double *JB = (double *)calloc((size_t)(11 * NG), sizeof(double));
    if (!P || !D || !JB) { free(P); free(D); free(JB); return -3; }
    orc_pad_mirror(nx, ny, nz, g, phi, P);
    orc_pad_mirror(nx, ny, nz, g, dmin, D);
    orc_pad_mirror(nx, ny, nz, g, jbn, JB);
    orc_pad_mirror(nx, ny, nz, g, jbn + 4 * N, JB + 4 * NG);
    orc_pad_mirror(nx, ny, nz, g, jbn + 8 * N, JB + 8 * NG);
    int rc = orc_ls_reinit_padded(gx, gy, gz, g, P, JB, D, dx, dy, dz, deltamin);
    if (!rc)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    phi[IDX(i, j, k)] =
                        P[(long)(i + g) + (long)gx * ((long)(j + g) + (long)gy * (k + g))];
    free(P); free(D); free(JB);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Dynamic Smagorinsky (Germano 1991, Lilly 1992 least squares).           */

/* Box test filter of ratio 2: weights (1/4,1/2,1/4) per direction with
 * clamped indices at the walls, applied x, then y, then z. */
/* Separable test filter: weights (ws, 1 - 2 ws, ws) along x, then y, then
 * z, indices clamped at the walls.  ws = 1/4 is the box filter of ratio 2. */
static void box3w(int nx, int ny, int nz, const double *in, double *out,
                  double *tmp, double ws)
{
    const double wc = 1.0 - 2.0 * ws;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int im = i > 0 ? i - 1 : 0, ip = i < nx - 1 ? i + 1 : nx - 1;
                out[IDX(i, j, k)] = ws * in[IDX(im, j, k)] + wc * in[IDX(i, j, k)]
                                  + ws * in[IDX(ip, j, k)];
            }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int jm = j > 0 ? j - 1 : 0, jp = j < ny - 1 ? j + 1 : ny - 1;
                tmp[IDX(i, j, k)] = ws * out[IDX(i, jm, k)] + wc * out[IDX(i, j, k)]
                                  + ws * out[IDX(i, jp, k)];
            }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                int km = k > 0 ? k - 1 : 0, kp = k < nz - 1 ? k + 1 : nz - 1;
                out[IDX(i, j, k)] = ws * tmp[IDX(i, j, km)] + wc * tmp[IDX(i, j, k)]
                                  + ws * tmp[IDX(i, j, kp)];
            }
}

/* nu_t = max(0, C * delta^2 * |S|), C = <L:M>/<M:M> with <.> the same box
 * filter, L_ij = hat(u_i u_j) - hat(u_i) hat(u_j),
 * M_ij = 2 delta^2 (hat(|S| S_ij) - 4 |hat S| hat S_ij), hat S = hat(S),
 * |S| = sqrt(2 S_ij S_ij); C = 0 where <M:M> = 0. */
int orc_dynsmag_cfg(int nx, int ny, int nz, const double *u, const double *gp,
                    const double *delta, double *nut, double ratio2, double ws,
                    double offw)
{
    long N = (long)nx * ny * nz;
    /* fields: S(6) |S| hatS(6) |hatS| SS-hat(6) uu-hat(6) uhat(3) LM MM tmp out */
    double *w = (double *)malloc(sizeof(double) * (size_t)N * 34);
    if (!w) return -3;
    double *S = w, *aS = w + 6 * N, *hS = w + 7 * N, *ahS = w + 13 * N,
           *hSS = w + 14 * N, *huu = w + 20 * N, *hu = w + 26 * N,
           *LM = w + 29 * N, *MM = w + 30 * N, *tmp = w + 31 * N,
           *t1 = w + 32 * N, *t2 = w + 33 * N;
    static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (long p = 0; p < N; p++) {
        double s2 = 0.0;
        for (int q = 0; q < 6; q++) {
            double s = 0.5 * (gp[(3 * ia[q] + ib[q]) * N + p] + gp[(3 * ib[q] + ia[q]) * N + p]);
            S[q * N + p] = s;
            s2 += (q < 3 ? 1.0 : offw) * s * s;
        }
        aS[p] = sqrt(2.0 * s2);
    }
    for (int q = 0; q < 6; q++) {
        box3w(nx, ny, nz, S + q * N, hS + q * N, tmp, ws);
        for (long p = 0; p < N; p++) t1[p] = aS[p] * S[q * N + p];
        box3w(nx, ny, nz, t1, hSS + q * N, tmp, ws);
        for (long p = 0; p < N; p++) t1[p] = u[ia[q] * N + p] * u[ib[q] * N + p];
        box3w(nx, ny, nz, t1, huu + q * N, tmp, ws);
    }
    for (int c = 0; c < 3; c++) box3w(nx, ny, nz, u + c * N, hu + c * N, tmp, ws);
    for (long p = 0; p < N; p++) {
        double s2 = 0.0;
        for (int q = 0; q < 6; q++)
            s2 += (q < 3 ? 1.0 : offw) * hS[q * N + p] * hS[q * N + p];
        ahS[p] = sqrt(2.0 * s2);
    }
    for (long p = 0; p < N; p++) {
        double lm = 0.0, mm = 0.0, d2 = delta[p] * delta[p];
        for (int q = 0; q < 6; q++) {
            double L = huu[q * N + p] - hu[ia[q] * N + p] * hu[ib[q] * N + p];
            double M = 2.0 * d2 * (hSS[q * N + p] - ratio2 * ahS[p] * hS[q * N + p]);
            double wq = q < 3 ? 1.0 : offw;
            lm += wq * L * M;
            mm += wq * M * M;
        }
        LM[p] = lm;
        MM[p] = mm;
    }
    box3w(nx, ny, nz, LM, t1, tmp, ws);
    box3w(nx, ny, nz, MM, t2, tmp, ws);
    for (long p = 0; p < N; p++) {
        double C = t2[p] > 0.0 ? t1[p] / t2[p] : 0.0;
        double v = C * delta[p] * delta[p] * aS[p];
        nut[p] = v > 0.0 ? v : 0.0;
    }
    free(w);
    return 0;
}

int orc_dynsmag(int nx, int ny, int nz, const double *u, const double *gp,
                const double *delta, double *nut)
{
    return orc_dynsmag_cfg(nx, ny, nz, u, gp, delta, nut, 4.0, 0.25, 2.0);
}


/* ---------------------------------------------------------------------- */
/* Adjoints.  The reference has no adjoint; these are the transposes of   */
/* the functions above, written from D = inv_h * A^{-1} B:                */
/*   D^T g = B^T A^{-T} g  (B including inv_h, as rhs_line forms it).     */
/* ---------------------------------------------------------------------- */

/* Columns of B: BT[7*l + k] = B[l+k-3][l], from rhs_line applied to the
 * unit vectors (n calls, once per 3-D call). */
static double *rhs_columns(int scheme, int n, double ih)
{
    double *BT = (double *)calloc((size_t)7 * n, sizeof(double));
    double *e = (double *)calloc((size_t)n, sizeof(double));
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    if (!BT || !e || !r) { free(BT); free(e); free(r); return NULL; }
    for (int l = 0; l < n; l++) {
        e[l] = 1.0;
        rhs_line(scheme, n, e, 1, ih, r);
        for (int k = 0; k < 7; k++) {
            int row = l + k - 3;
            if (row >= 0 && row < n) BT[7 * l + k] = r[row];
        }
        e[l] = 0.0;
    }
    free(e);
    free(r);
    return BT;
}

/* One line: solve A^T y = g (Thomas on the transposed tridiagonal: row l
 * is (c_{l-1}, b_l, a_{l+1})), then out = B^T y. */
static void deriv_line_adj(int n, const double *a, const double *b,
                           const double *c, const double *BT, const double *g,
                           long is, double *out, long os, double *y, double *gam)
{
    double bet = b[0];
    y[0] = g[0] / bet;
    for (int l = 1; l < n; l++) {
        double at = c[l - 1], ctm = a[l];   /* A^T[l][l-1], A^T[l-1][l] */
        gam[l] = ctm / bet;
        bet = b[l] - at * gam[l];
        y[l] = (g[(long)l * is] - at * y[l - 1]) / bet;
    }
    for (int l = n - 2; l >= 0; l--) y[l] = y[l] - gam[l + 1] * y[l + 1];
    for (int l = 0; l < n; l++) {
        double s = 0.0;
        for (int k = 0; k < 7; k++) {
            int t = l + k - 3;
            if (t >= 0 && t < n) s += BT[7 * l + k] * y[t];
        }
        out[(long)l * os] = s;
    }
}

int orc_deriv3d_adj(int scheme, int dir, int nx, int ny, int nz, const double *g,
                    double inv_h, double *out)
{
    int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    if (n < scheme_min_n(scheme)) return -1;
    double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)n);
    double *BT = rhs_columns(scheme, n, inv_h);
    if (!buf || !BT) { free(buf); free(BT); return -3; }
    double *a = buf, *b = buf + n, *c = buf + 2 * n, *y = buf + 3 * n, *gam = buf + 4 * n;
    orc_coeffs(scheme, n, a, b, c);
    long sx = 1, sy = nx, sz = (long)nx * ny;
    if (dir == 0) {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                deriv_line_adj(n, a, b, c, BT, g + IDX(0, j, k), sx, out + IDX(0, j, k), sx, y, gam);
    } else if (dir == 1) {
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                deriv_line_adj(n, a, b, c, BT, g + IDX(i, 0, k), sy, out + IDX(i, 0, k), sy, y, gam);
    } else {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                deriv_line_adj(n, a, b, c, BT, g + IDX(i, j, 0), sz, out + IDX(i, j, 0), sz, y, gam);
    }
    free(buf);
    free(BT);
    return 0;
}

int orc_deriv_vel_adj(int scheme, int nx, int ny, int nz, const double *dvel_bar,
                      const double *dtemp_bar, const double ih[3], double *u_bar,
                      double *temp_bar)
{
    long N = (long)nx * ny * nz;
    double *t = (double *)malloc(sizeof(double) * (size_t)N);
    if (!t) return -3;
    memset(u_bar, 0, sizeof(double) * 3 * (size_t)N);
    if (dtemp_bar) memset(temp_bar, 0, sizeof(double) * (size_t)N);
    for (int d = 0; d < 3; d++) {
        for (int c = 0; c < 3; c++) {
            if (orc_deriv3d_adj(scheme, d, nx, ny, nz, dvel_bar + (3 * d + c) * N, ih[d], t)) {
                free(t);
                return -1;
            }
            for (long p = 0; p < N; p++) u_bar[c * N + p] += t[p];
        }
        if (dtemp_bar) {
            if (orc_deriv3d_adj(scheme, d, nx, ny, nz, dtemp_bar + d * N, ih[d], t)) {
                free(t);
                return -1;
            }
            for (long p = 0; p < N; p++) temp_bar[p] += t[p];
        }
    }
    free(t);
    return 0;
}

int orc_divergencia_adj(int scheme, int nx, int ny, int nz, const double *rs_bar,
                        const double *rsv_bar, const double *jbn, double ihx,
                        double ihy, double ihz, double *flux_bar)
{
    long N = (long)nx * ny * nz;
    static const int mcomp[3] = {1, 5, 9};
    double ih[3] = {ihx, ihy, ihz};
    double *w = (double *)malloc(sizeof(double) * (size_t)N);
    double *d = (double *)malloc(sizeof(double) * (size_t)N);
    if (!w || !d) { free(w); free(d); return -3; }
    const double *ij = jbn + 10 * N, *det = jbn + 9 * N;
    int rc = 0;
    for (int q = 0; q < 6 && !rc; q++) {
        int dir = q % 3;
        const double *src = q < 3 ? rs_bar : rsv_bar;
        const double *met = jbn + (mcomp[dir] - 1) * N;
        for (long p = 0; p < N; p++) w[p] = ij[p] * src[p];
        rc = orc_deriv3d_adj(scheme, dir, nx, ny, nz, w, ih[dir], d);
        for (long p = 0; p < N && !rc; p++) flux_bar[q * N + p] = det[p] * met[p] * d[p];
    }
    free(w);
    free(d);
    return rc;
}

int orc_flujos_adj(int neq, int nx, int ny, int nz, const double *um,
                   const double *u, const double *vis, const double *jbn,
                   const double *dvel, const double *flux_bar, double *um_bar,
                   double *u_bar, double *pres_bar, double *vis_bar,
                   double *dvel_bar)
{
    long N = (long)nx * ny * nz;
    if (neq < 1 || neq > 5) return -1;
    for (long p = 0; p < N; p++) {
#define UM(m) um[((m) - 1) * N + p]
#define U(m) u[((m) - 1) * N + p]
#define J(c) jbn[((c) - 1) * N + p]
#define D(c) dvel[((c) - 1) * N + p]
#define UMB(m) um_bar[((m) - 1) * N + p]
#define UB(m) u_bar[((m) - 1) * N + p]
#define DB(c) dvel_bar[((c) - 1) * N + p]
        const double eb = flux_bar[p], fb = flux_bar[N + p], gb = flux_bar[2 * N + p];
        const double evb = flux_bar[3 * N + p], fvb = flux_bar[4 * N + p],
                     gvb = flux_bar[5 * N + p];
        const double mu = vis[p];
        switch (neq) {
        case 1: /* e,f,g = -um2,-um3,-um4 */
            UMB(2) -= eb; UMB(3) -= fb; UMB(4) -= gb;
            break;
        case 2: case 3: case 4: {
            int a = neq - 1; /* velocity component 1..3 */
            double cb[3] = {eb, fb, gb};
            for (int dd = 0; dd < 3; dd++) {
                UMB(2 + dd) -= cb[dd] * U(a);
                UB(a) -= cb[dd] * UM(2 + dd);
            }
            pres_bar[p] -= cb[a - 1];
            if (neq == 2) {
                /* ev = mu*2*J1*D1; fv = mu*(J1*D2+J5*D4); gv = mu*(J1*D3+J9*D7) */
                vis_bar[p] += evb * (2.0 * (J(1) * D(1))) + fvb * (J(1) * D(2) + J(5) * D(4))
                            + gvb * (J(1) * D(3) + J(9) * D(7));
                DB(1) += evb * mu * 2.0 * J(1);
                DB(2) += fvb * mu * J(1); DB(4) += fvb * mu * J(5);
                DB(3) += gvb * mu * J(1); DB(7) += gvb * mu * J(9);
            } else if (neq == 3) {
                /* ev = mu*(J1*D2+J5*D4); fv = mu*2*J5*D5; gv = mu*(J9*D8+J5*D6) */
                vis_bar[p] += evb * (J(1) * D(2) + J(5) * D(4)) + fvb * (2.0 * (J(5) * D(5)))
                            + gvb * (J(9) * D(8) + J(5) * D(6));
                DB(2) += evb * mu * J(1); DB(4) += evb * mu * J(5);
                DB(5) += fvb * mu * 2.0 * J(5);
                DB(8) += gvb * mu * J(9); DB(6) += gvb * mu * J(5);
            } else {
                /* ev = mu*(J1*D3+J9*D7); fv = mu*(J9*D8+J5*D6); gv = mu*2*J9*D9 */
                vis_bar[p] += evb * (J(1) * D(3) + J(9) * D(7)) + fvb * (J(9) * D(8) + J(5) * D(6))
                            + gvb * (2.0 * (J(9) * D(9)));
                DB(3) += evb * mu * J(1); DB(7) += evb * mu * J(9);
                DB(8) += fvb * mu * J(9); DB(6) += fvb * mu * J(5);
                DB(9) += gvb * mu * 2.0 * J(9);
            }
            break;
        }
        case 5: { /* e,f,g = -u1*um5, -u2*um5, -u3*um5 */
            double cb[3] = {eb, fb, gb};
            for (int dd = 0; dd < 3; dd++) {
                UB(1 + dd) -= cb[dd] * UM(5);
                UMB(5) -= cb[dd] * U(1 + dd);
            }
            break;
        }
        }
#undef UM
#undef U
#undef J
#undef D
#undef UMB
#undef UB
#undef DB
    }
    return 0;
}

int orc_flux_div_all_adj(int scheme, int nx, int ny, int nz, const double *rs_bar,
                         const double *rsv_bar, const double *um, const double *u,
                         const double *vis, const double *jbn, const double *dvel,
                         double ihx, double ihy, double ihz, double *um_bar,
                         double *u_bar, double *pres_bar, double *vis_bar,
                         double *dvel_bar)
{
    long N = (long)nx * ny * nz;
    double *fb = (double *)malloc(sizeof(double) * 6 * (size_t)N);
    if (!fb) return -3;
    memset(um_bar, 0, sizeof(double) * 5 * (size_t)N);
    memset(u_bar, 0, sizeof(double) * 3 * (size_t)N);
    memset(pres_bar, 0, sizeof(double) * (size_t)N);
    memset(vis_bar, 0, sizeof(double) * (size_t)N);
    memset(dvel_bar, 0, sizeof(double) * 9 * (size_t)N);
    int rc = 0;
    for (int m = 1; m <= 5 && !rc; m++) {
        rc = orc_divergencia_adj(scheme, nx, ny, nz, rs_bar + (m - 1) * N,
                                 rsv_bar + (m - 1) * N, jbn, ihx, ihy, ihz, fb);
        if (!rc)
            rc = orc_flujos_adj(m, nx, ny, nz, um, u, vis, jbn, dvel, fb, um_bar,
                                u_bar, pres_bar, vis_bar, dvel_bar);
    }
    free(fb);
    return rc;
}


/* ---------------------------------------------------------------------- */
/* Learned closure network (no reference counterpart).                    */
/* ---------------------------------------------------------------------- */
enum { MLP_H = 64, MLP_IN = 10 };
#define MLP_W1 0
#define MLP_B1 (MLP_W1 + MLP_H * MLP_IN)
#define MLP_W2 (MLP_B1 + MLP_H)
#define MLP_B2 (MLP_W2 + MLP_H * MLP_H)
#define MLP_W3 (MLP_B2 + MLP_H)
#define MLP_B3 (MLP_W3 + MLP_H * MLP_H)
#define MLP_W4 (MLP_B3 + MLP_H)
#define MLP_B4 (MLP_W4 + MLP_H)
#define MLP_NPARAM (MLP_B4 + 1)

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z)
{
    double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}
static double softplus(double z) { return z > 30.0 ? z : log1p(exp(z)); }

/* One point: pre-activations z1..z3 and activations h1..h3, returns z4. */
static double mlp_point(const float *P, const float *norm, const double *f,
                        double *z1, double *h1, double *z2, double *h2,
                        double *z3, double *h3)
{
    double x[MLP_IN];
    for (int k = 0; k < MLP_IN; k++) x[k] = (f[k] - norm[k]) * norm[MLP_IN + k];
    for (int o = 0; o < MLP_H; o++) {
        double s = P[MLP_B1 + o];
        for (int k = 0; k < MLP_IN; k++) s += P[MLP_W1 + o * MLP_IN + k] * x[k];
        z1[o] = s; h1[o] = silu(s);
    }
    for (int o = 0; o < MLP_H; o++) {
        double s = P[MLP_B2 + o];
        for (int k = 0; k < MLP_H; k++) s += P[MLP_W2 + o * MLP_H + k] * h1[k];
        z2[o] = s; h2[o] = silu(s);
    }
    for (int o = 0; o < MLP_H; o++) {
        double s = P[MLP_B3 + o];
        for (int k = 0; k < MLP_H; k++) s += P[MLP_W3 + o * MLP_H + k] * h2[k];
        z3[o] = s; h3[o] = silu(s);
    }
    double s = P[MLP_B4];
    for (int k = 0; k < MLP_H; k++) s += P[MLP_W4 + k] * h3[k];
    return s;
}

int orc_mlp_forward(const float *params, const float *norm, const double *feat,
                    long n, double *nu)
{
    double z1[MLP_H], h1[MLP_H], z2[MLP_H], h2[MLP_H], z3[MLP_H], h3[MLP_H];
    for (long p = 0; p < n; p++) {
        double z4 = mlp_point(params, norm, feat + p * MLP_IN, z1, h1, z2, h2, z3, h3);
        nu[p] = norm[2 * MLP_IN] * softplus(z4);
    }
    return 0;
}

int orc_mlp_grad(const float *params, const float *norm, const double *feat,
                 const double *dnu, long n, double *grad)
{
    const float *P = params;
    double z1[MLP_H], h1[MLP_H], z2[MLP_H], h2[MLP_H], z3[MLP_H], h3[MLP_H];
    double d1[MLP_H], d2[MLP_H], d3[MLP_H], x[MLP_IN];
    memset(grad, 0, sizeof(double) * MLP_NPARAM);
    for (long p = 0; p < n; p++) {
        const double *f = feat + p * MLP_IN;
        for (int k = 0; k < MLP_IN; k++) x[k] = (f[k] - norm[k]) * norm[MLP_IN + k];
        double z4 = mlp_point(P, norm, f, z1, h1, z2, h2, z3, h3);
        /* d softplus / dz = sigmoid(z) */
        double d4 = dnu[p] * norm[2 * MLP_IN] / (1.0 + exp(-z4));
        grad[MLP_B4] += d4;
        for (int k = 0; k < MLP_H; k++) {
            grad[MLP_W4 + k] += d4 * h3[k];
            d3[k] = d4 * P[MLP_W4 + k] * dsilu(z3[k]);
        }
        for (int o = 0; o < MLP_H; o++) {
            grad[MLP_B3 + o] += d3[o];
            for (int k = 0; k < MLP_H; k++) grad[MLP_W3 + o * MLP_H + k] += d3[o] * h2[k];
        }
        for (int k = 0; k < MLP_H; k++) {
            double s = 0.0;
            for (int o = 0; o < MLP_H; o++) s += P[MLP_W3 + o * MLP_H + k] * d3[o];
            d2[k] = s * dsilu(z2[k]);
        }
        for (int o = 0; o < MLP_H; o++) {
            grad[MLP_B2 + o] += d2[o];
            for (int k = 0; k < MLP_H; k++) grad[MLP_W2 + o * MLP_H + k] += d2[o] * h1[k];
        }
        for (int k = 0; k < MLP_H; k++) {
            double s = 0.0;
            for (int o = 0; o < MLP_H; o++) s += P[MLP_W2 + o * MLP_H + k] * d2[o];
            d1[k] = s * dsilu(z1[k]);
        }
        for (int o = 0; o < MLP_H; o++) {
            grad[MLP_B1 + o] += d1[o];
            for (int k = 0; k < MLP_IN; k++) grad[MLP_W1 + o * MLP_IN + k] += d1[o] * x[k];
        }
    }
    return 0;
}

int orc_mlp_features(int nx, int ny, int nz, const double *dvel, const double *jbn,
                     const double *delta, double *feat)
{
    long N = (long)nx * ny * nz;
    for (long p = 0; p < N; p++) {
        const double m[3] = {jbn[p], jbn[4 * N + p], jbn[8 * N + p]};
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                feat[p * MLP_IN + 3 * a + b] = m[b] * dvel[(3 * b + a) * N + p];
        feat[p * MLP_IN + 9] = delta[p];
    }
    return 0;
}


/* solver/rungek_ico.f90:87-91 (sub-step 1), 152-153 (2), 216-219 (3), in
 * the Fortran evaluation order. */
int orc_rk_update(int sub, long n, double *um1, const double *um0, const double *um,
                  const double *rs, const double *rsv, double dt)
{
    if (sub < 1 || sub > 3) return -1;
    for (long p = 0; p < n; p++) {
        double r = rs[p] + rsv[p];
        if (sub == 1) um1[p] = um0[p] + dt * r;
        else if (sub == 2) um1[p] = 0.75 * um0[p] + 0.25 * um[p] + 0.25 * dt * r;
        else um1[p] = (1.0 / 3.0) * um0[p] + (2.0 / 3.0) * um[p] + (2.0 / 3.0) * dt * r;
    }
    return 0;
}


/* ---- RUNGEK glue around the operator ------------------------------------ */

#define UM1(i, j, k, m) um1[(i) + (long)nx * ((j) + (long)ny * ((k) + (long)nz * (m)))]

/* frontera_z_val, frontera_y_val, frontera_x_val (cited in the header),
 * statement by statement per boundary column so overlapping edges see the
 * same order as the Fortran loops. */
int orc_frontera_val(int nx, int ny, int nz, double *um1)
{
    if (nx < 3 || ny < 3 || nz < 3) return -3;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            UM1(i, j, 0, 0) = UM1(i, j, 1, 0);
            UM1(i, j, 0, 1) = UM1(i, j, 1, 1);
            UM1(i, j, 0, 2) = UM1(i, j, 1, 2);
            UM1(i, j, 0, 3) = 0.0;
            UM1(i, j, 0, 4) = UM1(i, j, 1, 4);
            UM1(i, j, nz - 1, 0) = 1.0;
            UM1(i, j, nz - 2, 0) = 1.0;
            UM1(i, j, nz - 1, 1) = UM1(i, j, nz - 3, 1);
            UM1(i, j, nz - 2, 1) = UM1(i, j, nz - 3, 1);
            UM1(i, j, nz - 1, 2) = UM1(i, j, nz - 3, 2);
            UM1(i, j, nz - 2, 2) = UM1(i, j, nz - 3, 2);
            UM1(i, j, nz - 1, 3) = 0.0;
            UM1(i, j, nz - 2, 3) = 0.0;
            UM1(i, j, nz - 1, 4) = UM1(i, j, nz - 3, 4);
            UM1(i, j, nz - 2, 4) = UM1(i, j, nz - 3, 4);
        }
    for (int k = 0; k < nz; k++)
        for (int i = 0; i < nx; i++) {
            UM1(i, 0, k, 0) = UM1(i, 1, k, 0);
            UM1(i, ny - 1, k, 0) = UM1(i, ny - 2, k, 0);
            UM1(i, 0, k, 1) = 0.0;
            UM1(i, ny - 1, k, 1) = 0.0;
            UM1(i, 0, k, 2) = 0.0;
            UM1(i, ny - 1, k, 2) = 0.0;
            UM1(i, 0, k, 3) = 0.0;
            UM1(i, ny - 1, k, 3) = 0.0;
            UM1(i, 0, k, 4) = UM1(i, 1, k, 4);
            UM1(i, ny - 1, k, 4) = UM1(i, ny - 2, k, 4);
        }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
            UM1(0, j, k, 1) = 0.0;
            UM1(0, j, k, 2) = 0.0;
            UM1(0, j, k, 3) = 0.0;
            UM1(0, j, k, 4) = UM1(1, j, k, 4);
            UM1(0, j, k, 0) = UM1(1, j, k, 0);
            UM1(nx - 1, j, k, 1) = 0.0;
            UM1(nx - 1, j, k, 2) = 0.0;
            UM1(nx - 1, j, k, 3) = 0.0;
            UM1(nx - 1, j, k, 4) = UM1(nx - 2, j, k, 4);
            UM1(nx - 1, j, k, 0) = UM1(nx - 2, j, k, 0);
        }
    return 0;
}
#undef UM1

int orc_varc_to_varnc(long n, const double *um1, double mach, double *u, double *pres,
                      double *temp, double *conc)
{
    for (long p = 0; p < n; p++) {
        if (conc) conc[p] = um1[4 * n + p];
        temp[p] = um1[4 * n + p];
        pres[p] = um1[p] / (mach * mach);
        for (int c = 0; c < 3; c++) u[c * n + p] = um1[(1 + c) * n + p] / um1[p];
    }
    return 0;
}

int orc_varnc_to_varc(long n, const double *u, const double *pres, const double *temp,
                      double mach, double *um0, double *pres0)
{
    for (long p = 0; p < n; p++) {
        um0[p] = pres[p] * mach * mach;
        um0[4 * n + p] = temp[p];
        for (int c = 0; c < 3; c++) um0[(1 + c) * n + p] = u[c * n + p] * um0[p];
        if (pres0) pres0[p] = pres[p];
    }
    return 0;
}

int orc_rk_clip(long n, double *um1, const double *mask, double mach, double *u, double *pres,
                double *temp, double *conc)
{
    orc_varc_to_varnc(n, um1, mach, u, pres, temp, conc);
    for (long p = 0; p < n; p++)
        for (int c = 0; c < 3; c++) {
            double v = u[c * n + p];
            v = v < 6.0 ? v : 6.0;     /* min(6, u) */
            v = v > -6.0 ? v : -6.0;   /* max(-6, u) */
            u[c * n + p] = v;
            um1[(1 + c) * n + p] = v * um1[p] * mask[p];
        }
    return orc_varc_to_varnc(n, um1, mach, u, pres, temp, conc);
}

int orc_rk_rebuild(long n, double *um1, const double *u, const double *pres, double mach)
{
    for (long p = 0; p < n; p++) {
        um1[p] = pres[p] * mach * mach;
        for (int c = 0; c < 3; c++) um1[(1 + c) * n + p] = u[c * n + p] * um1[p];
    }
    return 0;
}


/* ---- compact filter ------------------------------------------------------ */

int orc_filter_coeffs(int periodic, int n, double cfilt, double *a, double *b, double *c)
{
    if (n < 2) return -3;
    for (int l = 0; l < n; l++) {
        b[l] = 1.0;
        /* INIFILTRO sets cfilt2..cfilt5 to zero on both branches */
        const int interior = l + 1 >= 6 && l + 1 <= n - 5;
        a[l] = c[l] = (periodic || interior) ? cfilt : 0.0;
    }
    return 0;
}

/* filtronper: 11-point explicit RHS on rows 6..n-5 (1-based), identity on
 * the five rows at each end, then the Thomas algorithm in the reference's
 * form (division by the pivot). */
int orc_filtronper(int n, double *u, const double *a, const double *b, const double *c,
                   double cfilt)
{
    if (n < 2) return -3;
    const double arsf = (193. + 126. * cfilt) / 256.;
    const double brsf = (105. + 302. * cfilt) / 256.;
    const double crsf = (-15. + 30. * cfilt) / 64.;
    const double drsf = (45. - 90. * cfilt) / 512.;
    const double ersf = (-5. + 10. * cfilt) / 256.;
    const double frsf = (1. - 2 * cfilt) / 512.;
    double *r = malloc(sizeof(double) * (size_t)n);
    double *gam = malloc(sizeof(double) * (size_t)(n + 1));
    if (!r || !gam) { free(r); free(gam); return -5; }
    for (int l = 0; l < n; l++) {
        if (l + 1 >= 6 && l + 1 <= n - 5)
            r[l] = 0.5 * (arsf * (u[l] + u[l]) + brsf * (u[l + 1] + u[l - 1]) +
                          crsf * (u[l + 2] + u[l - 2]) + drsf * (u[l + 3] + u[l - 3]) +
                          ersf * (u[l + 4] + u[l - 4]) + frsf * (u[l + 5] + u[l - 5]));
        else
            r[l] = u[l];
    }
    double bet = b[0];
    u[0] = r[0] / bet;
    for (int l = 1; l < n; l++) {
        gam[l] = c[l - 1] / bet;
        bet = b[l] - a[l] * gam[l];
        u[l] = (r[l] - a[l] * u[l - 1]) / bet;
    }
    for (int l = n - 2; l >= 0; l--) u[l] = u[l] - gam[l + 1] * u[l + 1];
    free(r);
    free(gam);
    return 0;
}

/* one directional pass of FILTRADO over field m */
static int filter_dir(int dir, int nx, int ny, int nz, double *f, const double *j10,
                      const double *j11, double cfilt)
{
    const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
    const long sl = dir == 0 ? 1 : (dir == 1 ? nx : (long)nx * ny);
    double *a = malloc(sizeof(double) * (size_t)n), *b = malloc(sizeof(double) * (size_t)n);
    double *c = malloc(sizeof(double) * (size_t)n), *line = malloc(sizeof(double) * (size_t)n);
    if (!a || !b || !c || !line) { free(a); free(b); free(c); free(line); return -5; }
    orc_filter_coeffs(dir == 2, n, cfilt, a, b, c);
    const int n1 = dir == 0 ? ny : nx, n2 = dir == 2 ? ny : nz;
    int rc = 0;
    for (int o2 = 0; o2 < n2 && !rc; o2++)
        for (int o1 = 0; o1 < n1 && !rc; o1++) {
            long base;
            if (dir == 0) base = (long)nx * (o1 + (long)ny * o2);
            else if (dir == 1) base = o1 + (long)nx * ny * o2;
            else base = o1 + (long)nx * o2;
            for (int l = 0; l < n; l++) line[l] = f[base + l * sl] * j10[base + l * sl];
            rc = orc_filtronper(n, line, a, b, c, cfilt);
            for (int l = 0; l < n; l++) f[base + l * sl] = line[l] * j11[base + l * sl];
        }
    free(a); free(b); free(c); free(line);
    return rc;
}

int orc_filtrado(int itf, int nx, int ny, int nz, double *um1, const double *jbn, double cfilt,
                 double mach, double *u, double *pres, double *temp)
{
    /* direction order of each itf (herramientas_incomp2.f90:245-392) */
    static const int order[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                    {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    if (itf < 1 || itf > 6) return -1;
    const long N = (long)nx * ny * nz;
    for (int m = 0; m < 5; m++)
        for (int k = 0; k < 3; k++) {
            int rc = filter_dir(order[itf - 1][k], nx, ny, nz, um1 + m * N, jbn + 9 * N,
                                jbn + 10 * N, cfilt);
            if (rc) return rc;
        }
    for (long p = 0; p < N; p++) {
        for (int c = 0; c < 3; c++) u[c * N + p] = um1[(1 + c) * N + p] / um1[p];
        temp[p] = um1[4 * N + p];
        pres[p] = um1[p] / (mach * mach);
    }
    return 0;
}
