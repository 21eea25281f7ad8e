/*
 * ORACLE — test infrastructure only.  Accessors around the reference's own
 * table code, compiled together with the unmodified reference sources
 * potential.c, errs.c and cell.c into _ref/libmdref.so (see Makefile).
 * This file is ours; it only reads the reference's struct potential and
 * instantiates its header-inline potential_eval (potential_eval.h:67-104).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fptype.h"
#include "potential.h"
#include "potential_eval.h"

int ref_potential_data(const struct potential *p, int *n, double *ab, double *alpha, double *c, int csize) {
    int k;
    if (p == NULL) return -1;
    *n = p->n;
    ab[0] = p->a;
    ab[1] = p->b;
    for (k = 0; k < 4; k++) alpha[k] = p->alpha[k];
    if ((p->n + 1) * potential_chunk > csize) return -2;
    for (k = 0; k < (p->n + 1) * potential_chunk; k++) c[k] = p->c[k];
    return 0;
}

void ref_potential_eval(struct potential *p, double r2, double *e, double *f) {
    FPTYPE ee, ff;
    potential_eval(p, r2, &ee, &ff);
    *e = ee;
    *f = ff;
}

/* Lennard-Jones + reaction field (the host library's potential_create_LJ126_RF
   formula), built by the reference's potential_init extension point. */
static double rf_A, rf_B, rf_q, rf_ke, rf_krf, rf_crf;
static double rf_f(double r) {
    return potential_LJ126(r, rf_A, rf_B) + rf_ke * rf_q * (1.0 / r + rf_krf * r * r - rf_crf);
}
static double rf_fp(double r) {
    return potential_LJ126_p(r, rf_A, rf_B) + rf_ke * rf_q * (-1.0 / (r * r) + 2.0 * rf_krf * r);
}
static double rf_f6p(double r) {
    double r2 = r * r, r4 = r2 * r2, r7 = r * r2 * r4;
    return potential_LJ126_6p(r, rf_A, rf_B) + 720.0 * rf_ke * rf_q / r7;
}

struct potential *ref_potential_create_LJ126_RF(double a, double b, double A, double B, double q, double ke,
                                                double eps_rf, double tol) {
    struct potential *p;
    rf_A = A;
    rf_B = B;
    rf_q = q;
    rf_ke = ke;
    rf_krf = (eps_rf <= 0.0) ? 1.0 / (2.0 * b * b * b) : (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * b * b * b);
    rf_crf = 1.0 / b + rf_krf * b * b;
    if ((p = (struct potential *)malloc(sizeof(struct potential))) == NULL) return NULL;
    if (potential_init(p, rf_f, rf_fp, rf_f6p, a, b, tol) < 0) {
        free(p);
        return NULL;
    }
    return p;
}
