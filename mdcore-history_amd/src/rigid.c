/*
 * Rigid distance constraints: engine_rigid_add (engine_rigid.c:78-230),
 * engine_rigid_eval (engine_rigid.c:352-583, one thread: the groups in
 * order), rigid_eval_shake (rigid.c:81-277) and rigid_eval_pshake
 * (rigid.c:279-523).
 *
 * Both solvers run after the advance on positions x(t+dt) and velocities
 * v(t+dt/2).  The constraint directions are taken from the positions
 * before the drift, x - dt v.  Each iteration computes the residues
 * d^2 - |r_ij|^2 and moves the particles along the old directions, weighted
 * by the masses, until every |residue| < tol_rigid.  The new positions then
 * set the velocities, v += (x_new - x) / dt.  P-SHAKE couples the
 * constraints through a matrix a (float), refined on the first
 * engine_pshake_steps steps or when more than rigid_pshake_refine
 * iterations were needed, and also adds the constraint forces to f.
 * Particles of a group may sit in neighbouring cells: positions are taken
 * relative to the first particle's cell with a shift of at most one cell
 * per axis.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdcore.h"

#define eerror_(id) errs_register((id), engine_err_msg[-(id)], __LINE__, __func__, __FILE__)

extern const char *engine_err_msg[];

int engine_rigid_add(struct engine *e, int pid, int pjd, double d) {
    struct rigid *r, *q;
    int rid, rjd, k, j, ind, jnd;
    if (e == NULL) return eerror_(engine_err_null);
    if (pid < 0 || pjd < 0 || pid == pjd || pid >= e->s.size_parts || pjd >= e->s.size_parts)
        return eerror_(engine_err_range);
    if (e->part2rigid == NULL || e->part2rigid_size < e->s.size_parts) {
        int *tmp = (int *)malloc(sizeof(int) * e->s.size_parts);
        if (tmp == NULL) return eerror_(engine_err_malloc);
        for (k = 0; k < e->s.size_parts; k++) tmp[k] = k < e->part2rigid_size ? e->part2rigid[k] : -1;
        free(e->part2rigid);
        e->part2rigid = tmp;
        e->part2rigid_size = e->s.size_parts;
    }
    rid = e->part2rigid[pid];
    rjd = e->part2rigid[pjd];
    if (rid < 0 && rjd < 0) {
        /* a new group of two */
        if (e->nr_rigids == e->rigids_size) {
            int size = e->rigids_size < 16 ? 16 : (int)(e->rigids_size * 1.414);
            struct rigid *tmp = (struct rigid *)malloc(sizeof(struct rigid) * size);
            if (tmp == NULL) return eerror_(engine_err_malloc);
            if (e->nr_rigids) memcpy(tmp, e->rigids, sizeof(struct rigid) * e->nr_rigids);
            free(e->rigids);
            e->rigids = tmp;
            e->rigids_size = size;
        }
        r = &e->rigids[e->nr_rigids];
        r->nr_parts = 2;
        r->parts[0] = pid;
        r->parts[1] = pjd;
        r->nr_constr = 1;
        r->constr[0].i = 0;
        r->constr[0].j = 1;
        r->constr[0].d2 = d * d;
        e->part2rigid[pid] = e->part2rigid[pjd] = e->nr_rigids;
        e->nr_rigids += 1;
    } else if (rid >= 0 && rjd >= 0 && rid != rjd) {
        /* two groups: append rjd's particles and constraints to rid, add
           the new constraint, and fill rjd's slot with the last group */
        r = &e->rigids[rid];
        q = &e->rigids[rjd];
        if (r->nr_parts + q->nr_parts > rigid_maxparts || r->nr_constr + q->nr_constr + 1 > rigid_maxconstr)
            return eerror_(engine_err_range);
        for (ind = 0; r->parts[ind] != pid; ind++)
            ;
        for (jnd = 0; q->parts[jnd] != pjd; jnd++)
            ;
        for (j = 0; j < q->nr_constr; j++) {
            r->constr[r->nr_constr + j] = q->constr[j];
            r->constr[r->nr_constr + j].i += r->nr_parts;
            r->constr[r->nr_constr + j].j += r->nr_parts;
        }
        for (j = 0; j < q->nr_parts; j++) {
            r->parts[r->nr_parts + j] = q->parts[j];
            e->part2rigid[q->parts[j]] = rid;
        }
        r->constr[r->nr_constr + q->nr_constr].i = ind;
        r->constr[r->nr_constr + q->nr_constr].j = r->nr_parts + jnd;
        r->constr[r->nr_constr + q->nr_constr].d2 = d * d;
        r->nr_constr += q->nr_constr + 1;
        r->nr_parts += q->nr_parts;
        e->nr_rigids -= 1;
        if (rjd < e->nr_rigids) {
            e->rigids[rjd] = e->rigids[e->nr_rigids];
            for (j = 0; j < e->rigids[rjd].nr_parts; j++) e->part2rigid[e->rigids[rjd].parts[j]] = rjd;
            if (rid == e->nr_rigids) rid = rjd; /* the merged group itself moved */
        }
        r = &e->rigids[rid];
    } else {
        /* one group holds one or both particles */
        if (rid < 0) rid = rjd;
        r = &e->rigids[rid];
        ind = jnd = -1;
        for (j = 0; j < r->nr_parts; j++) {
            if (r->parts[j] == pid)
                ind = j;
            else if (r->parts[j] == pjd)
                jnd = j;
        }
        if ((ind < 0 || jnd < 0) && r->nr_parts >= rigid_maxparts) return eerror_(engine_err_range);
        if (r->nr_constr >= rigid_maxconstr) return eerror_(engine_err_range);
        if (ind < 0) {
            ind = r->nr_parts++;
            r->parts[ind] = pid;
            e->part2rigid[pid] = rid;
        } else if (jnd < 0) {
            jnd = r->nr_parts++;
            r->parts[jnd] = pjd;
            e->part2rigid[pjd] = rid;
        }
        r->constr[r->nr_constr].i = ind;
        r->constr[r->nr_constr].j = jnd;
        r->constr[r->nr_constr].d2 = d * d;
        r->nr_constr += 1;
    }
    e->nr_constr += 1;
    /* the coupling matrix starts as the identity */
    memset(r->a, 0, sizeof(float) * r->nr_constr * r->nr_constr);
    for (k = 0; k < r->nr_constr; k++) r->a[k + k * r->nr_constr] = 1.0f;
    return engine_err_ok;
}

/* The group's particles, masses and positions relative to the first
   particle's cell; 0 if a particle is missing. */
static int rigid_load(const struct rigid *r, struct engine *e, struct part **p, int *shift, double *m, double *xp) {
    const struct space *s = &e->s;
    const struct cell *c0 = NULL;
    int k, j;
    for (k = 0; k < r->nr_parts; k++) {
        const struct cell *c;
        if ((p[k] = s->partlist[r->parts[k]]) == NULL) return 0;
        c = s->celllist[r->parts[k]];
        if (k == 0) c0 = c;
        m[k] = e->types[p[k]->type].mass;
        for (j = 0; j < 3; j++) {
            int sh = 0;
            if (c != c0) {
                sh = c->loc[j] - c0->loc[j];
                if (sh > 1)
                    sh = -1;
                else if (sh < -1)
                    sh = 1;
            }
            shift[3 * k + j] = sh;
            xp[3 * k + j] = (c != c0) ? p[k]->x[j] + s->h[j] * sh : p[k]->x[j];
        }
    }
    return 1;
}

/* Write the solved positions back; velocities (and, for P-SHAKE, forces)
   take the correction. */
static void rigid_store(const struct rigid *r, struct engine *e, struct part **p, const int *shift,
                        const double *m, const double *xp, double idt, int forces) {
    int k, j;
    for (k = 0; k < r->nr_parts; k++)
        for (j = 0; j < 3; j++) {
            const double xn = (e->s.celllist[r->parts[k]] != e->s.celllist[r->parts[0]])
                                  ? xp[3 * k + j] - e->s.h[j] * shift[3 * k + j]
                                  : xp[3 * k + j];
            if (forces) p[k]->f[j] += idt * idt * (xn - p[k]->x[j]) * m[k];
            p[k]->v[j] += idt * (xn - p[k]->x[j]);
            p[k]->x[j] = xn;
        }
}

int rigid_eval_shake(struct rigid *rs, int N, struct engine *e) {
    struct part *p[rigid_maxparts];
    int shift[3 * rigid_maxparts], rid, iter, k, j, pid, pjd;
    double m[rigid_maxparts], xp[3 * rigid_maxparts], xo[3 * rigid_maxparts];
    double vc[3 * rigid_maxconstr], wvc[3 * rigid_maxconstr], res[rigid_maxconstr];
    double dt, idt, tol, lambda, w, max_res;
    if (rs == NULL || e == NULL) return rigid_err_null;
    dt = e->dt;
    idt = 1.0 / dt;
    tol = e->tol_rigid;
    for (rid = 0; rid < N; rid++) {
        struct rigid *r = &rs[rid];
        if (!rigid_load(r, e, p, shift, m, xp)) continue;
        for (k = 0; k < r->nr_parts; k++)
            for (j = 0; j < 3; j++) xo[3 * k + j] = xp[3 * k + j] - dt * p[k]->v[j];
        for (k = 0; k < r->nr_constr; k++) {
            pid = r->constr[k].i;
            pjd = r->constr[k].j;
            for (j = 0; j < 3; j++) {
                vc[3 * k + j] = xo[3 * pid + j] - xo[3 * pjd + j];
                wvc[3 * k + j] = 1.0 / (m[pid] + m[pjd]) * vc[3 * k + j];
            }
        }
        for (iter = 0; iter < rigid_maxiter; iter++) {
            for (max_res = 0.0, k = 0; k < r->nr_constr; k++) {
                pid = r->constr[k].i;
                pjd = r->constr[k].j;
                res[k] = r->constr[k].d2;
                for (j = 0; j < 3; j++) res[k] -= (xp[3 * pid + j] - xp[3 * pjd + j]) * (xp[3 * pid + j] - xp[3 * pjd + j]);
                if (fabs(res[k]) > max_res) max_res = fabs(res[k]);
            }
            if (max_res < tol) break;
            for (k = 0; k < r->nr_constr; k++) {
                pid = r->constr[k].i;
                pjd = r->constr[k].j;
                lambda = 0.5 * res[k] /
                         ((xp[3 * pid] - xp[3 * pjd]) * vc[3 * k] + (xp[3 * pid + 1] - xp[3 * pjd + 1]) * vc[3 * k + 1] +
                          (xp[3 * pid + 2] - xp[3 * pjd + 2]) * vc[3 * k + 2]);
                for (j = 0; j < 3; j++) {
                    w = lambda * wvc[3 * k + j];
                    xp[3 * pid + j] += w * m[pjd];
                    xp[3 * pjd + j] -= w * m[pid];
                }
            }
        }
        rigid_store(r, e, p, shift, m, xp, idt, 0);
    }
    return rigid_err_ok;
}

int rigid_eval_pshake(struct rigid *rs, int N, struct engine *e, int a_update) {
    struct part *p[rigid_maxparts];
    int shift[3 * rigid_maxparts], rid, iter, k, j, i, pid, pjd, np, nc;
    double m[rigid_maxparts], xp[3 * rigid_maxparts], xo[3 * rigid_maxparts], res[rigid_maxconstr];
    double vc[3 * rigid_maxconstr * rigid_maxparts];
    double dv[3], dt, idt, tol, lambda, max_res;
    if (rs == NULL || e == NULL) return rigid_err_null;
    dt = e->dt;
    idt = 1.0 / dt;
    tol = e->tol_rigid;
    for (rid = 0; rid < N; rid++) {
        struct rigid *r = &rs[rid];
        np = r->nr_parts;
        nc = r->nr_constr;
        if (!rigid_load(r, e, p, shift, m, xp)) continue;
        for (k = 0; k < np; k++)
            for (j = 0; j < 3; j++) xo[3 * k + j] = xp[3 * k + j] - dt * p[k]->v[j];
        /* vc[i] = sum_k a[k][i] * (mass-weighted gradient of constraint k) */
        memset(vc, 0, sizeof(double) * 3 * nc * np);
        for (k = 0; k < nc; k++) {
            pid = r->constr[k].i;
            pjd = r->constr[k].j;
            for (j = 0; j < 3; j++) dv[j] = (xo[3 * pid + j] - xo[3 * pjd + j]) / (m[pid] + m[pjd]);
            for (i = 0; i < nc; i++)
                if (r->a[k * nc + i] != 0.0f)
                    for (j = 0; j < 3; j++) {
                        vc[i * np * 3 + pid * 3 + j] += r->a[k * nc + i] * dv[j] * m[pjd];
                        vc[i * np * 3 + pjd * 3 + j] += -r->a[k * nc + i] * dv[j] * m[pid];
                    }
        }
        for (iter = 0; iter < rigid_maxiter; iter++) {
            for (max_res = 0.0, k = 0; k < nc; k++) {
                const double *g = &vc[k * 3 * np];
                pid = r->constr[k].i;
                pjd = r->constr[k].j;
                res[k] = r->constr[k].d2;
                for (j = 0; j < 3; j++) res[k] -= (xp[3 * pid + j] - xp[3 * pjd + j]) * (xp[3 * pid + j] - xp[3 * pjd + j]);
                if (fabs(res[k]) > max_res) max_res = fabs(res[k]);
                lambda = 0.5 * res[k] /
                         ((xp[3 * pid] - xp[3 * pjd]) * (g[3 * pid] - g[3 * pjd]) +
                          (xp[3 * pid + 1] - xp[3 * pjd + 1]) * (g[3 * pid + 1] - g[3 * pjd + 1]) +
                          (xp[3 * pid + 2] - xp[3 * pjd + 2]) * (g[3 * pid + 2] - g[3 * pjd + 2]));
                for (j = 0; j < 3 * np; j++) xp[j] += lambda * g[j];
            }
            if (max_res < tol) break;
        }
        if (nc > 1 && (a_update || iter > rigid_pshake_refine)) {
            /* refine the coupling: a <- a * a_new, with a_new[i][j] =
               -(dx_i . g_j) / (dx_i . g_i) off the diagonal, 1 on it,
               scaled down when its largest entry exceeds the limit */
            float a_new[rigid_maxconstr * rigid_maxconstr], tmp[rigid_maxconstr * rigid_maxconstr];
            float w, dx[3], max_alpha = 0.0f;
            for (i = 0; i < nc; i++) {
                const double *gi = &vc[i * 3 * np];
                pid = r->constr[i].i;
                pjd = r->constr[i].j;
                for (j = 0; j < 3; j++) dx[j] = xp[3 * pid + j] - xp[3 * pjd + j];
                w = -1.0 / (dx[0] * (gi[pid * 3] - gi[pjd * 3]) + dx[1] * (gi[pid * 3 + 1] - gi[pjd * 3 + 1]) +
                            dx[2] * (gi[pid * 3 + 2] - gi[pjd * 3 + 2]));
                for (j = i + 1; j < nc; j++) {
                    const double *gj = &vc[j * 3 * np];
                    a_new[i + j * nc] = w * (dx[0] * (gj[pid * 3] - gj[pjd * 3]) + dx[1] * (gj[pid * 3 + 1] - gj[pjd * 3 + 1]) +
                                             dx[2] * (gj[pid * 3 + 2] - gj[pjd * 3 + 2]));
                    a_new[j + i * nc] = a_new[i + j * nc];
                    max_alpha = fmaxf(max_alpha, fabsf(a_new[i + j * nc]));
                }
                a_new[i + i * nc] = 1.0f;
            }
            if (max_alpha > rigid_pshake_maxalpha) {
                w = 0.1 / max_alpha;
                for (i = 0; i < nc * nc; i++) a_new[i] *= w;
                for (i = 0; i < nc; i++) a_new[i * nc + i] = 1.0f;
            }
            memset(tmp, 0, sizeof(float) * nc * nc);
            for (j = 0; j < nc; j++)
                for (k = 0; k < nc; k++)
                    for (i = 0; i < nc; i++) tmp[j * nc + i] += r->a[k * nc + i] * a_new[j * nc + k];
            memcpy(r->a, tmp, sizeof(float) * nc * nc);
        }
        rigid_store(r, e, p, shift, m, xp, idt, 1);
    }
    return rigid_err_ok;
}

int engine_rigid_eval(struct engine *e) {
    if (e == NULL) return eerror_(engine_err_null);
    if (e->nr_rigids == 0) return engine_err_ok;
    if (e->flags & engine_flag_shake) {
        if (rigid_eval_shake(e->rigids, e->nr_rigids, e) < 0) return eerror_(engine_err_rigid);
    } else if (rigid_eval_pshake(e->rigids, e->nr_rigids, e, e->time < engine_pshake_steps) < 0)
        return eerror_(engine_err_rigid);
    return engine_err_ok;
}

int engine_getrigid(struct engine *e, int rid, int *parts, int *ij, double *d2) {
    const struct rigid *r;
    int k;
    if (e == NULL) return eerror_(engine_err_null);
    if (rid == -1) return e->nr_rigids;
    if (rid < 0 || rid >= e->nr_rigids) return eerror_(engine_err_range);
    r = &e->rigids[rid];
    if (parts) {
        parts[0] = r->nr_parts;
        for (k = 0; k < r->nr_parts; k++) parts[1 + k] = r->parts[k];
    }
    for (k = 0; k < r->nr_constr; k++) {
        if (ij) {
            ij[2 * k] = r->constr[k].i;
            ij[2 * k + 1] = r->constr[k].j;
        }
        if (d2) d2[k] = r->constr[k].d2;
    }
    return r->nr_constr;
}
