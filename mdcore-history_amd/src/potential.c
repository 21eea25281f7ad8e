/*
 * Piecewise-quintic interpolation tables (host side, FP64).
 *
 * The table format and the construction steps follow the reference's
 * potential.c so that the tables are bit-identical to the ones it builds:
 *   potential_getalpha      potential.c:1537-1600  (golden-section search for
 *                                                    the node-density exponent)
 *   potential_init          potential.c:963-1194   (interval-count search,
 *                                                    Newton node placement,
 *                                                    index transform, left
 *                                                    extension chunk)
 *   potential_getcoeffs     potential.c:1401-1520  (Chebyshev-corrected Hermite
 *                                                    quintic per interval)
 *   fit of the node slopes  potential.c:1207-1374  (tridiagonal least squares)
 *   potential_eval          potential_eval.h:67-104
 * One deliberate difference: the per-node value/slope arrays hold n+1
 * entries (the reference declares n and writes n+1, potential.c:1404; with
 * gcc's default stack layout that overrun makes every table build fail).
 * Every floating-point expression keeps the reference's operation order,
 * because a different order changes the last bits of the coefficients.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdcore.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

int potential_err = potential_err_ok;

static FPTYPE c_null[potential_chunk] = {0, 0, 0, 0, 0, 0, 0, 0};
struct potential potential_null = {{0, 0, 0, 0}, c_null, 0.0, DBL_MAX, potential_flag_none, 1};

static const char *potential_err_msg[6] = {
    "Nothing bad happened.",
    "An unexpected NULL pointer was encountered.",
    "A call to malloc failed, probably due to insufficient memory.",
    "The requested value was out of bounds.",
    "Not yet implemented.",
    "Maximum number of intervals reached before tolerance satisfied."};

#define perror_(id) \
    (potential_err = errs_register((id), potential_err_msg[-(id)], __LINE__, __func__, __FILE__))

/* ------------------------------------------------------------ analytic forms */

double potential_switch(double r, double A, double B) {
    if (r < A) return 1.0;
    if (r > B) return 0.0;
    double B2 = B * B, A2 = A * A, r2 = r * r;
    double B2mr2 = B2 - r2, B2mA2 = B2 - A2;
    return B2mr2 * B2mr2 * (B2 + 2 * r2 - 3 * A2) / (B2mA2 * B2mA2 * B2mA2);
}

double potential_switch_p(double r, double A, double B) {
    if (A < r && r < B) {
        double B2 = B * B, A2 = A * A, r2 = r * r;
        double B2mr2 = B2 - r2, B2mA2 = B2 - A2;
        double r2_p = 2 * r, B2mr2_p = -r2_p;
        return (2 * B2mr2_p * B2mr2 * (B2 + 2 * r2 - 3 * A2) + B2mr2 * B2mr2 * 2 * r2_p) /
               (B2mA2 * B2mA2 * B2mA2);
    }
    return 0.0;
}

double potential_LJ126(double r, double A, double B) {
    double ir = 1.0 / r, ir2 = ir * ir, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
    return (A * ir12 - B * ir6);
}

double potential_LJ126_p(double r, double A, double B) {
    double ir = 1.0 / r, ir2 = ir * ir, ir4 = ir2 * ir2, ir12 = ir4 * ir4 * ir4;
    return 6.0 * ir * (-2.0 * A * ir12 + B * ir4 * ir2);
}

double potential_LJ126_6p(double r, double A, double B) {
    double r2 = r * r, ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
    return 10080.0 * ir12 * (884.0 * A * ir6 - 33.0 * B);
}

double potential_Coulomb(double r) { return potential_escale / r; }

double potential_Coulomb_p(double r) { return -potential_escale / (r * r); }

double potential_Coulomb_6p(double r) {
    double r2 = r * r, r4 = r2 * r2, r7 = r * r2 * r4;
    return 720.0 * potential_escale / r7;
}

double potential_Ewald(double r, double kappa) { return potential_escale * erfc(kappa * r) / r; }

double potential_Ewald_p(double r, double kappa) {
    double r2 = r * r, ir = 1.0 / r, ir2 = ir * ir;
    const double isqrtpi = 0.56418958354775628695;
    return potential_escale * (-2.0 * exp(-kappa * kappa * r2) * kappa * ir * isqrtpi - erfc(kappa * r) * ir2);
}

double potential_Ewald_6p(double r, double kappa) {
    double r2 = r * r, ir2 = 1.0 / r2, r4 = r2 * r2, ir4 = ir2 * ir2, ir6 = ir2 * ir4;
    double k2 = kappa * kappa;
    const double isqrtpi = 0.56418958354775628695;
    double t6 = erfc(kappa * r), t23 = exp(-k2 * r2);
    return potential_escale *
           (720.0 * t6 / r * ir6 +
            (1440.0 * ir6 + (960.0 * ir4 + (384.0 * ir2 + (144.0 + (-128.0 * r2 + 64.0 * k2 * r4) * k2) * k2) * k2) * k2) *
                kappa * isqrtpi * t23);
}

/* ------------------------------------------------------------ node exponent */

double potential_getalpha(double (*f6p)(double), double a, double b) {
    double xi[potential_N], fx[potential_N];
    double alpha[4], fa[4], maxf = 0.0, t;
    const double golden = 2.0 / (1 + sqrt(5));
    int i, j;

    for (i = 0; i < potential_N; i++) {
        xi[i] = ((double)i + 1) / (potential_N + 1);
        fx[i] = f6p(a + (b - a) * xi[i]);
        maxf = fmax(maxf, fabs(fx[i]));
    }
    if (maxf == 0.0) return 1.0;

    /* Golden-section bracket on [0,2] of max_x |(alpha + 2(1-alpha)x)^-6 f6(x)|. */
    alpha[0] = 0;
    alpha[3] = 2;
    alpha[1] = alpha[3] - 2 * golden;
    alpha[2] = alpha[0] + 2 * golden;
    for (i = 0; i < 4; i++) {
        fa[i] = 0.0;
        for (j = 0; j < potential_N; j++) {
            t = fabs(pow(alpha[i] + 2 * (1 - alpha[i]) * xi[j], -6) * fx[j]);
            if (t > fa[i]) fa[i] = t;
        }
    }
    while (alpha[3] - alpha[0] > 1.0e-4) {
        if (fa[1] < fa[2]) {
            alpha[3] = alpha[2]; fa[3] = fa[2];
            alpha[2] = alpha[1]; fa[2] = fa[1];
            alpha[1] = alpha[3] - (alpha[3] - alpha[0]) * golden;
            i = 1;
        } else {
            alpha[0] = alpha[1]; fa[0] = fa[1];
            alpha[1] = alpha[2]; fa[1] = fa[2];
            alpha[2] = alpha[0] + (alpha[3] - alpha[0]) * golden;
            i = 2;
        }
        fa[i] = 0.0;
        for (j = 0; j < potential_N; j++) {
            t = fabs(pow(alpha[i] + 2 * (1 - alpha[i]) * xi[j], -6) * fx[j]);
            if (t > fa[i]) fa[i] = t;
        }
    }
    return (alpha[0] + alpha[3]) / 2;
}

/* ------------------------------------------------------------ node slopes */

/* Weighted quadrature nodes/weights shared by both slope fits: Chebyshev
   points cos(k pi/(N-1)) with weight 1/sqrt(1-x^2), end weights zero. */
static double *quad_w = NULL, *quad_x = NULL;

static int quad_init(void) {
    int k;
    if (quad_w != NULL) return 0;
    quad_w = (double *)malloc(sizeof(double) * potential_N);
    quad_x = (double *)malloc(sizeof(double) * potential_N);
    if (quad_w == NULL || quad_x == NULL) return -1;
    for (k = 1; k < potential_N - 1; k++) {
        quad_x[k] = cos(k * M_PI / (potential_N - 1));
        quad_w[k] = 1.0 / sqrt(1.0 - quad_x[k] * quad_x[k]);
    }
    quad_x[0] = 1.0;
    quad_x[potential_N - 1] = -1.0;
    quad_w[0] = 0.0;
    quad_w[potential_N - 1] = 0.0;
    return 0;
}

/* Least-squares node slopes fp[0..n].  With fixed ends (fpa, fpb given) the
   first and last rows are identities (potential.c:1294-1374), otherwise
   they are the free-end rows (potential.c:1207-1290). */
static int fit_node_slopes(double (*f)(double), int fixed, double fpa, double fpb, int n, FPTYPE *x,
                           double *fp) {
    /* Projections of the Hermite basis onto the slope cardinal functions on
       [-1,1] (the constants of the reference). */
    static const double cwl1[4] = {0.25, -0.25, -0.25, 0.25};
    static const double cwr1[4] = {-0.25, -0.25, 0.25, 0.25};
    const double wl0wl1 = 0.1125317885884428;
    const double wl1wl1 = 0.03215579530433858;
    const double wl0wr1 = -0.04823369227661384;
    const double wl1wr1 = -0.02143719641629633;
    const double wr0wr1 = -0.1125317885884429;
    const double wr1wr1 = 0.03215579530433859;
    const double wl1wr0 = 0.04823369227661384;
    int i, k;
    double m, h, eff;
    double *fx, *hx, *d0, *d1, *d2, *rhs, *vl, *vr;

    if (quad_init() < 0) return perror_(potential_err_malloc);
    fx = (double *)malloc(sizeof(double) * 8 * (n + 1));
    if (fx == NULL) return perror_(potential_err_malloc);
    hx = fx + (n + 1);
    d0 = hx + (n + 1);
    d1 = d0 + (n + 1);
    d2 = d1 + (n + 1);
    rhs = d2 + (n + 1);
    vl = rhs + (n + 1);
    vr = vl + (n + 1);

    for (i = 0; i <= n; i++) fx[i] = f(x[i]);
    for (i = 0; i < n; i++) hx[i] = x[i + 1] - x[i];
    for (i = 0; i < n; i++) {
        vl[i] = 0.0;
        vr[i] = 0.0;
        m = 0.5 * (x[i] + x[i + 1]);
        h = 0.5 * (x[i + 1] - x[i]);
        for (k = 1; k < potential_N - 1; k++) {
            const double t = quad_x[k];
            eff = f(m + h * t);
            vl[i] += quad_w[k] * (eff * (cwl1[0] + t * (cwl1[1] + t * (cwl1[2] + t * cwl1[3]))));
            vr[i] += quad_w[k] * (eff * (cwr1[0] + t * (cwr1[1] + t * (cwr1[2] + t * cwr1[3]))));
        }
        vl[i] /= potential_N - 2;
        vr[i] /= potential_N - 2;
    }

    if (fixed) {
        d1[0] = 1.0;
        d2[0] = 0.0;
        rhs[0] = fpa;
    } else {
        d1[0] = wl1wl1 * hx[0];
        d2[0] = wl1wr1 * hx[0];
        rhs[0] = 2 * (vl[0] - fx[0] * wl0wl1 - fx[1] * wl1wr0);
    }
    for (i = 1; i < n; i++) {
        d0[i] = wl1wr1 * hx[i - 1];
        d1[i] = wr1wr1 * hx[i - 1] + wl1wl1 * hx[i];
        d2[i] = wl1wr1 * hx[i];
        rhs[i] = 2 * (vr[i - 1] - fx[i - 1] * wl0wr1 - fx[i] * wr0wr1) +
                 2 * (vl[i] - fx[i] * wl0wl1 - fx[i + 1] * wl1wr0);
    }
    if (fixed) {
        d0[n] = 0.0;
        d1[n] = 1.0;
        rhs[n] = fpb;
    } else {
        d0[n] = wl1wr1 * hx[n - 1];
        d1[n] = wr1wr1 * hx[n - 1];
        rhs[n] = 2 * (vr[n - 1] - fx[n - 1] * wl0wr1 - fx[n] * wr0wr1);
    }

    /* Thomas elimination, then back substitution. */
    for (i = 1; i <= n; i++) {
        m = d0[i] / d1[i - 1];
        d1[i] = d1[i] - m * d2[i - 1];
        rhs[i] = rhs[i] - m * rhs[i - 1];
    }
    fp[n] = rhs[n] / d1[n];
    for (i = n - 1; i >= 0; i--) fp[i] = (rhs[i] - d2[i] * fp[i + 1]) / d1[i];

    free(fx);
    return potential_err_ok;
}

/* ------------------------------------------------------------ coefficients */

static FPTYPE *cheb_cos = NULL; /* cos(j k pi / N), k = 0..6 rows */

int potential_getcoeffs(double (*f)(double), double (*fp)(double), FPTYPE *xi, int n, FPTYPE *c,
                        FPTYPE *err) {
    int i, j, k, ind;
    double phi[7], cee[6], fa, fb, dfa, dfb;
    double h, m, w, e, err_loc, maxf, x;
    double fx[potential_N];
    double *fix, *fpx;

    if (f == NULL || xi == NULL || err == NULL) return perror_(potential_err_null);
    if (cheb_cos == NULL) {
        if ((cheb_cos = (FPTYPE *)malloc(sizeof(FPTYPE) * 7 * potential_N)) == NULL)
            return perror_(potential_err_malloc);
        for (k = 0; k < 7; k++)
            for (j = 0; j < potential_N; j++) cheb_cos[k * potential_N + j] = cos(j * k * M_PI / potential_N);
    }

    /* Node values and slopes: n+1 entries each. */
    if ((fix = (double *)malloc(sizeof(double) * 2 * (n + 1))) == NULL) return perror_(potential_err_malloc);
    fpx = fix + (n + 1);
    for (k = 0; k <= n; k++) fix[k] = f(xi[k]);
    if (fp == NULL) {
        if (fit_node_slopes(f, 0, 0.0, 0.0, n, xi, fpx) < 0) {
            free(fix);
            return perror_(potential_err);
        }
    } else {
        if (fit_node_slopes(f, 1, fp(xi[0]), fp(xi[n]), n, xi, fpx) < 0) {
            free(fix);
            return perror_(potential_err);
        }
    }

    *err = 0.0;
    for (i = 0; i < n; i++) {
        ind = i * (potential_degree + 3);
        m = (xi[i] + xi[i + 1]) / 2;
        h = (xi[i + 1] - xi[i]) / 2;
        fa = fix[i];
        fb = fix[i + 1];
        dfa = fpx[i] * h;
        dfb = fpx[i + 1] * h;

        /* Chebyshev coefficients phi_0..6 of f on the interval. */
        for (k = 0; k < potential_N; k++) fx[k] = f(m + h * cos(k * M_PI / potential_N));
        for (j = 0; j < 7; j++) {
            phi[j] = (fa + (1 - 2 * (j % 2)) * fb) / 2;
            for (k = 1; k < potential_N; k++) phi[j] += fx[k] * cheb_cos[j * potential_N + k];
            phi[j] *= 2.0 / potential_N;
        }

        /* Cubic Hermite part in the Chebyshev basis ... */
        cee[0] = (4 * (fa + fb) + dfa - dfb) / 4;
        cee[1] = -(9 * (fa - fb) + dfa + dfb) / 16;
        cee[2] = (dfb - dfa) / 8;
        cee[3] = (fa - fb + dfa + dfb) / 16;
        cee[4] = 0.0;
        cee[5] = 0.0;
        /* ... plus the T4 and T5 corrections that keep the end values and
           slopes and best match phi_4, phi_5. */
        w = (6 * (cee[0] - phi[0]) - 4 * (cee[2] - phi[2]) - phi[4]) / (36 + 16 + 1);
        cee[0] += -6 * w;
        cee[2] += 4 * w;
        cee[4] = -w;
        w = (2 * (cee[1] - phi[1]) - 3 * (cee[3] - phi[3]) - phi[5]) / (4 + 9 + 1);
        cee[1] += -2 * w;
        cee[3] += 3 * w;
        cee[5] = -w;

        /* Monomials on [-1,1]: chunk = {m, 1/h, x^5, x^4, x^3, x^2, x^1, x^0}. */
        c[ind + 7] = cee[0] / 2 - cee[2] + cee[4];
        c[ind + 6] = cee[1] - 3 * cee[3] + 5 * cee[5];
        c[ind + 5] = 2 * cee[2] - 8 * cee[4];
        c[ind + 4] = 4 * cee[3] - 20 * cee[5];
        c[ind + 3] = 8 * cee[4];
        c[ind + 2] = 16 * cee[5];
        c[ind + 1] = 1.0 / h;
        c[ind] = m;

        /* Max error on the Chebyshev points, relative to max(|f|, 1). */
        maxf = 0.0;
        err_loc = 0.0;
        for (k = 1; k < potential_N; k++) {
            maxf = fmax(fabs(fx[k]), maxf);
            x = cheb_cos[potential_N + k];
            e = fabs(fx[k] - c[ind + 7] -
                     x * (c[ind + 6] + x * (c[ind + 5] + x * (c[ind + 4] + x * (c[ind + 3] + x * c[ind + 2])))));
            err_loc = fmax(e, err_loc);
        }
        err_loc /= fmax(maxf, 1.0);
        *err = fmax(err_loc, *err);
    }
    free(fix);
    return potential_err_ok;
}

/* ------------------------------------------------------------ table build */

/* Place the n+1 nodes so that the index transform maps node i to i/n:
   Newton on i - n*(a0 + x(a1 + x a2)) = 0 (potential.c:1008-1016). */
static void place_nodes(const struct potential *p, FPTYPE a, FPTYPE b, int n, FPTYPE *xi) {
    const FPTYPE mtol = 10 * DBL_EPSILON;
    int i;
    double e;
    xi[0] = a;
    xi[n] = b;
    for (i = 1; i < n; i++) {
        xi[i] = a + (b - a) * i / n;
        while (1) {
            e = i - n * (p->alpha[0] + xi[i] * (p->alpha[1] + xi[i] * p->alpha[2]));
            xi[i] += e / (n * (p->alpha[1] + 2 * xi[i] * p->alpha[2]));
            if (fabs(e) < n * mtol) break;
        }
    }
}

/* Build a trial table with n intervals into c[(n+1)*chunk] (chunk 0 left
   free for the extension) and return its error estimate. */
static int trial_table(const struct potential *p, double (*f)(double), double (*fp)(double), FPTYPE a,
                       FPTYPE b, int n, FPTYPE **c_out, FPTYPE *err) {
    FPTYPE *xi, *c;
    if ((xi = (FPTYPE *)malloc(sizeof(FPTYPE) * (n + 1))) == NULL) return perror_(potential_err_malloc);
    if (posix_memalign((void **)&c, potential_align, sizeof(FPTYPE) * (n + 1) * potential_chunk) != 0) {
        free(xi);
        return perror_(potential_err_malloc);
    }
    place_nodes(p, a, b, n, xi);
    if (potential_getcoeffs(f, fp, xi, n, &c[potential_chunk], err) < 0) {
        free(xi);
        free(c);
        return perror_(potential_err);
    }
    free(xi);
    *c_out = c;
    return potential_err_ok;
}

/* Scale the transform to the table index and pull the right edge b inside
   interval n (potential.c:1150-1165). */
static void finish_transform(struct potential *p, FPTYPE a, FPTYPE b) {
    double err_b;
    FPTYPE bee, aye, fbee;
    p->alpha[0] *= p->n;
    p->alpha[1] *= p->n;
    p->alpha[2] *= p->n;
    p->alpha[0] += 1.0;
    bee = copysign(sqrt(b * b), b);
    aye = copysign(sqrt(a * a), a);
    fbee = fabs(bee);
    err_b = (p->n + 1) - (p->alpha[0] + bee * (p->alpha[1] + bee * p->alpha[2])) -
            DBL_EPSILON * (fabs(p->alpha[0]) + fbee * (fabs(p->alpha[1]) + fbee * fabs(p->alpha[2])));
    p->alpha[0] += (err_b * aye) / (aye - bee);
    p->alpha[1] += (-err_b) / (aye - bee);
}

/* Chunk 0: the quadratic Taylor extension of interval 1 at its left end a,
   written in the variable x = (r - a)/a (potential.c:1167-1185). */
static void fill_left_extension(struct potential *p, FPTYPE a) {
    double coeffs[potential_degree], eff[potential_degree];
    int i, k;
    p->c[0] = a;
    p->c[1] = (a == 0.0) ? 1.0 : 1.0 / a;
    for (k = 0; k < potential_degree; k++) {
        coeffs[k] = p->c[2 * potential_chunk - 1 - k];
        eff[k] = 0.0;
    }
    for (i = 0; i < potential_degree; i++)
        for (k = potential_degree - 1; k >= i; k--) {
            eff[i] = coeffs[k] + (-1.0) * eff[i];
            coeffs[k] *= (k - i) * p->c[potential_chunk + 1] * a;
        }
    p->c[potential_chunk - 1] = eff[0];
    p->c[potential_chunk - 2] = eff[1];
    p->c[potential_chunk - 3] = 0.5 * eff[2];
    for (k = 3; k <= potential_degree; k++) p->c[potential_chunk - 1 - k] = 0.0;
}

int potential_init(struct potential *p, double (*f)(double), double (*fp)(double), double (*f6p)(double),
                   FPTYPE a, FPTYPE b, FPTYPE tol) {
    double alpha, w;
    int l = potential_ivalsa, r = potential_ivalsb, m;
    FPTYPE err_l, err_r, err_m;
    FPTYPE *c_l = NULL, *c_r = NULL, *c_m = NULL;

    if (p == NULL || f == NULL) return perror_(potential_err_null);
    p->a = a;
    p->b = b;
    p->flags = potential_flag_none;

    alpha = (f6p != NULL) ? potential_getalpha(f6p, a, b) : 1.0;

    /* Quadratic index transform: node i sits where a0 + x(a1 + x a2) = i/n. */
    w = 1.0 / (a - b);
    w *= w;
    p->alpha[0] = a * a * w - alpha * b * a * w;
    p->alpha[1] = -2 * a * w + alpha * (a + b) * w;
    p->alpha[2] = w - alpha * w;
    p->alpha[3] = 0.0;

    /* One interval may already do. */
    if (trial_table(p, f, fp, a, b, l, &c_l, &err_l) < 0) return perror_(potential_err);
    if (err_l < tol) {
        p->n = l;
        p->c = c_l;
        finish_transform(p, a, b);
        fill_left_extension(p, a);
        return potential_err_ok;
    }

    /* Double until the error bound holds ... */
    while (1) {
        if (trial_table(p, f, fp, a, b, r, &c_r, &err_r) < 0) {
            free(c_l);
            return perror_(potential_err);
        }
        if (err_r < tol) break;
        if (2 * r > potential_ivalsmax) {
            free(c_l);
            free(c_r);
            return perror_(potential_err_ivalsmax);
        }
        l = r;
        err_l = err_r;
        free(c_l);
        c_l = c_r;
        r *= 2;
    }

    /* ... then bisect for the smallest n that meets it. */
    while (r - l > 1) {
        m = 0.5 * (r + l);
        if (trial_table(p, f, fp, a, b, m, &c_m, &err_m) < 0) {
            free(c_l);
            free(c_r);
            return perror_(potential_err);
        }
        if (err_m > tol) {
            l = m;
            err_l = err_m;
            free(c_l);
            c_l = c_m;
        } else {
            r = m;
            err_r = err_m;
            free(c_r);
            c_r = c_m;
        }
    }
    (void)err_l;
    (void)err_r;

    p->n = r;
    p->c = c_r;
    finish_transform(p, a, b);
    fill_left_extension(p, a);
    free(c_l);
    return potential_err_ok;
}

void potential_clear(struct potential *p) {
    if (p == NULL) return;
    p->flags = potential_flag_none;
    free(p->c);
    p->c = NULL;
}

void potential_free(struct potential *p) {
    if (p == NULL || p == &potential_null) return;
    potential_clear(p);
    free(p);
}

int potential_getdata(const struct potential *p, int *n, double *ab, double *alpha, double *c, int csize) {
    int k;
    if (p == NULL) return perror_(potential_err_null);
    if (n) *n = p->n;
    if (ab) {
        ab[0] = p->a;
        ab[1] = p->b;
    }
    if (alpha)
        for (k = 0; k < 4; k++) alpha[k] = p->alpha[k];
    if (c) {
        if (csize < (p->n + 1) * potential_chunk) return perror_(potential_err_bounds);
        memcpy(c, p->c, sizeof(double) * (p->n + 1) * potential_chunk);
    }
    return potential_err_ok;
}

/* ------------------------------------------------------------ evaluation */

void potential_eval(const struct potential *p, FPTYPE r2, FPTYPE *e, FPTYPE *f) {
    int ind, k;
    FPTYPE x, ee, eff, r;
    const FPTYPE *c;
    r = sqrt(r2);
    ind = fmax(0.0, p->alpha[0] + r * (p->alpha[1] + r * p->alpha[2]));
    c = &p->c[ind * potential_chunk];
    x = (r - c[0]) * c[1];
    ee = c[2] * x + c[3];
    eff = c[2];
    for (k = 4; k < potential_chunk; k++) {
        eff = eff * x + ee;
        ee = ee * x + c[k];
    }
    *e = ee;
    *f = eff * c[1] / r;
}

/* Energy and dE/dr at the distances r[0..n) (a vector helper for callers
   that sample a table, e.g. to build training data). */
int potential_eval_r(const struct potential *p, const double *r, long long n, double *e, double *dedr) {
    long long k;
    FPTYPE ee, ff;
    if (p == NULL || r == NULL || e == NULL || dedr == NULL) return perror_(potential_err_null);
    for (k = 0; k < n; k++) {
        potential_eval(p, r[k] * r[k], &ee, &ff);
        e[k] = ee;
        dedr[k] = ff * r[k];
    }
    return potential_err_ok;
}

/* ------------------------------------------------------------ constructors */

/* The reference passes the parameters to potential_init's callbacks through
   file-scope variables (potential.c:301-914); so does this file.  These
   constructors are therefore not re-entrant, as in the reference. */
static double pc_A, pc_B, pc_q, pc_kappa, pc_s, pc_b, pc_K, pc_r0, pc_ke, pc_krf, pc_crf;

static struct potential *alloc_potential(void) {
    struct potential *p;
    if (posix_memalign((void **)&p, 16, sizeof(struct potential)) != 0) {
        perror_(potential_err_malloc);
        return NULL;
    }
    memset(p, 0, sizeof(*p));
    return p;
}

static struct potential *build(double (*f)(double), double (*fp)(double), double (*f6p)(double), double a,
                               double b, double tol, unsigned int flags) {
    struct potential *p = alloc_potential();
    if (p == NULL) return NULL;
    if (potential_init(p, f, fp, f6p, a, b, tol) < 0) {
        free(p);
        return NULL;
    }
    p->flags = flags;
    return p;
}

static double lj_f(double r) { return potential_LJ126(r, pc_A, pc_B); }
static double lj_fp(double r) { return potential_LJ126_p(r, pc_A, pc_B); }
static double lj_f6p(double r) { return potential_LJ126_6p(r, pc_A, pc_B); }

struct potential *potential_create_LJ126(double a, double b, double A, double B, double tol) {
    pc_A = A;
    pc_B = B;
    return build(lj_f, lj_fp, lj_f6p, a, b, tol, potential_flag_LJ126);
}

static double ljs_f(double r) { return potential_LJ126(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b); }
static double ljs_fp(double r) {
    return potential_LJ126_p(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b) +
           potential_LJ126(r, pc_A, pc_B) * potential_switch_p(r, pc_s, pc_b);
}
static double ljs_f6p(double r) { return potential_LJ126_6p(r, pc_A, pc_B); }

struct potential *potential_create_LJ126_switch(double a, double b, double A, double B, double s, double tol) {
    pc_A = A;
    pc_B = B;
    pc_s = s;
    pc_b = b;
    return build(ljs_f, ljs_fp, ljs_f6p, a, b, tol, potential_flag_LJ126);
}

static double lje_f(double r) { return potential_LJ126(r, pc_A, pc_B) + pc_q * potential_Ewald(r, pc_kappa); }
static double lje_fp(double r) {
    return potential_LJ126_p(r, pc_A, pc_B) + pc_q * potential_Ewald_p(r, pc_kappa);
}
static double lje_f6p(double r) {
    return potential_LJ126_6p(r, pc_A, pc_B) + pc_q * potential_Ewald_6p(r, pc_kappa);
}

struct potential *potential_create_LJ126_Ewald(double a, double b, double A, double B, double q, double kappa,
                                               double tol) {
    pc_A = A;
    pc_B = B;
    pc_q = q;
    pc_kappa = kappa;
    return build(lje_f, lje_fp, lje_f6p, a, b, tol, potential_flag_LJ126 | potential_flag_Ewald);
}

static double ljes_f(double r) {
    return potential_LJ126(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b) + pc_q * potential_Ewald(r, pc_kappa);
}
static double ljes_fp(double r) {
    return potential_LJ126_p(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b) +
           potential_LJ126(r, pc_A, pc_B) * potential_switch_p(r, pc_s, pc_b) +
           pc_q * potential_Ewald_p(r, pc_kappa);
}

struct potential *potential_create_LJ126_Ewald_switch(double a, double b, double A, double B, double q,
                                                      double kappa, double s, double tol) {
    pc_A = A;
    pc_B = B;
    pc_q = q;
    pc_kappa = kappa;
    pc_s = s;
    pc_b = b;
    return build(ljes_f, ljes_fp, lje_f6p, a, b, tol, potential_flag_LJ126 | potential_flag_Ewald);
}

static double ew_f(double r) { return pc_q * potential_Ewald(r, pc_kappa); }
static double ew_fp(double r) { return pc_q * potential_Ewald_p(r, pc_kappa); }
static double ew_f6p(double r) { return pc_q * potential_Ewald_6p(r, pc_kappa); }

struct potential *potential_create_Ewald(double a, double b, double q, double kappa, double tol) {
    pc_q = q;
    pc_kappa = kappa;
    return build(ew_f, ew_fp, ew_f6p, a, b, tol, potential_flag_Ewald);
}

static double co_f(double r) { return potential_escale * pc_q * (1.0 / r - 1.0 / pc_b); }
static double co_fp(double r) { return -potential_escale * pc_q / (r * r); }
static double co_f6p(double r) {
    double r2 = r * r, r4 = r2 * r2, r7 = r * r2 * r4;
    return 720.0 * potential_escale * pc_q / r7;
}

struct potential *potential_create_Coulomb(double a, double b, double q, double tol) {
    pc_q = q;
    pc_b = b;
    return build(co_f, co_fp, co_f6p, a, b, tol, potential_flag_Coulomb);
}

static double ljc_f(double r) {
    return potential_LJ126(r, pc_A, pc_B) + potential_escale * pc_q * (1.0 / r - 1.0 / pc_b);
}
static double ljc_fp(double r) { return potential_LJ126_p(r, pc_A, pc_B) - potential_escale * pc_q / (r * r); }
static double ljc_f6p(double r) {
    double r2 = r * r, r4 = r2 * r2, r7 = r * r2 * r4;
    return potential_LJ126_6p(r, pc_A, pc_B) + 720.0 * potential_escale * pc_q / r7;
}

struct potential *potential_create_LJ126_Coulomb(double a, double b, double A, double B, double q, double tol) {
    pc_q = q;
    pc_b = b;
    pc_A = A;
    pc_B = B;
    return build(ljc_f, ljc_fp, ljc_f6p, a, b, tol, potential_flag_LJ126 | potential_flag_Coulomb);
}

static double ha_f(double r) { return pc_K * (r - pc_r0) * (r - pc_r0); }
static double ha_f6p(double r) {
    (void)r;
    return 0;
}

struct potential *potential_create_harmonic(double a, double b, double K, double r0, double tol) {
    pc_K = K;
    pc_r0 = r0;
    return build(ha_f, NULL, ha_f6p, a, b, tol, potential_flag_none);
}

/* Reaction field (new, see mdcore.h). */
static double rf_f(double r) {
    return potential_LJ126(r, pc_A, pc_B) + pc_ke * pc_q * (1.0 / r + pc_krf * r * r - pc_crf);
}
static double rf_fp(double r) {
    return potential_LJ126_p(r, pc_A, pc_B) + pc_ke * pc_q * (-1.0 / (r * r) + 2.0 * pc_krf * r);
}
static double rf_f6p(double r) {
    double r2 = r * r, r4 = r2 * r2, r7 = r * r2 * r4;
    return potential_LJ126_6p(r, pc_A, pc_B) + 720.0 * pc_ke * pc_q / r7;
}

static double rfs_f(double r) {
    return potential_LJ126(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b) +
           pc_ke * pc_q * (1.0 / r + pc_krf * r * r - pc_crf);
}
static double rfs_fp(double r) {
    return potential_LJ126_p(r, pc_A, pc_B) * potential_switch(r, pc_s, pc_b) +
           potential_LJ126(r, pc_A, pc_B) * potential_switch_p(r, pc_s, pc_b) +
           pc_ke * pc_q * (-1.0 / (r * r) + 2.0 * pc_krf * r);
}

static void rf_constants(double b, double eps_rf) {
    pc_krf = (eps_rf <= 0.0) ? 1.0 / (2.0 * b * b * b) : (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * b * b * b);
    pc_crf = 1.0 / b + pc_krf * b * b;
}

struct potential *potential_create_LJ126_switch_RF(double a, double b, double A, double B, double s, double q,
                                                   double ke, double eps_rf, double tol) {
    pc_A = A;
    pc_B = B;
    pc_s = s;
    pc_b = b;
    pc_q = q;
    pc_ke = ke;
    rf_constants(b, eps_rf);
    return build(rfs_f, rfs_fp, rf_f6p, a, b, tol,
                 ((A != 0.0 || B != 0.0) ? potential_flag_LJ126 : 0) | potential_flag_Coulomb);
}

struct potential *potential_create_LJ126_RF(double a, double b, double A, double B, double q, double ke,
                                            double eps_rf, double tol) {
    pc_A = A;
    pc_B = B;
    pc_q = q;
    pc_ke = ke;
    rf_constants(b, eps_rf);
    return build(rf_f, rf_fp, rf_f6p, a, b, tol,
                 ((A != 0.0 || B != 0.0) ? potential_flag_LJ126 : 0) | potential_flag_Coulomb);
}
