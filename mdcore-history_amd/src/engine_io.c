/*
 * Structure and parameter readers (engine_io.c of the reference) and the
 * topology lists they fill.
 *
 *   engine_read_psf    engine_io.c:1234-1618: PSF header, NATOM (one type
 *                      per atom type and charge, named type + charge text),
 *                      NBOND / NTHETA / NPHI / NIMP into the topology lists,
 *                      then PDB ATOM records (Angstrom -> nm, wrapped).
 *   engine_read_cpf    engine_io.c:630-1232: CHARMM parameter file, sections
 *                      BOND, ANGL, DIHE, IMPR, NONB ... HBOND.
 *   engine_read_xplor  engine_io.c:82-628: XPLOR keyword lines (BOND, ANGL,
 *                      DIHE, IMPR, NONB), in any order.
 *
 * Both parameter readers end in the same construction of the nonbonded
 * tables (nb_tables): for every type pair with a Lennard-Jones well or, with
 * electrostatics, two charges, eps = sqrt(eps_i eps_j), r0 = rmin_i + rmin_j,
 * A = eps r0^12, B = 2 eps r0^6, q = q_i q_j; the inner radius is where the
 * potential reaches engine_maxKcutoff times its value at r0 (bisection down
 * to 1e-5 nm from [r0/2, r0]); pairs whose two types repeat the parameters
 * of earlier types share that earlier pair's table.  This library evaluates
 * no bonded terms, so bonded parameters are read past; with rigidH, bonds
 * between types whose first type starts with 'H' become rigid constraints
 * (d = r0 in nm) and leave the bond list, as in the reference.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdcore.h"
#include "reader.h"

extern const char *engine_err_msg[];
#define eerror_(id) errs_register((id), engine_err_msg[-(id)], __LINE__, __func__, __FILE__)

/* ---------------------------------------------------------------- topology */

#define GROW(arr, nr, size, type)                                                   \
    do {                                                                          \
        if ((nr) == (size)) {                                                     \
            int ns_ = (size) < 64 ? 64 : (int)((size) * 1.414);                   \
            type *t_ = (type *)malloc(sizeof(type) * ns_);                        \
            if (t_ == NULL) return eerror_(engine_err_malloc);                    \
            if (nr) memcpy(t_, (arr), sizeof(type) * (nr));                       \
            free(arr);                                                            \
            (arr) = t_;                                                           \
            (size) = ns_;                                                         \
        }                                                                         \
    } while (0)

int engine_bond_add(struct engine *e, int i, int j) {
    if (e == NULL) return eerror_(engine_err_null);
    GROW(e->bonds, e->nr_bonds, e->bonds_size, struct bond);
    e->bonds[e->nr_bonds].i = i;
    e->bonds[e->nr_bonds].j = j;
    e->bonds[e->nr_bonds].pid = -1;
    e->nr_bonds += 1;
    return engine_err_ok;
}

int engine_angle_add(struct engine *e, int i, int j, int k, int pid) {
    if (e == NULL) return eerror_(engine_err_null);
    GROW(e->angles, e->nr_angles, e->angles_size, struct angle);
    e->angles[e->nr_angles].i = i;
    e->angles[e->nr_angles].j = j;
    e->angles[e->nr_angles].k = k;
    e->angles[e->nr_angles].pid = pid;
    e->nr_angles += 1;
    return engine_err_ok;
}

int engine_dihedral_add(struct engine *e, int i, int j, int k, int l, int pid) {
    if (e == NULL) return eerror_(engine_err_null);
    GROW(e->dihedrals, e->nr_dihedrals, e->dihedrals_size, struct dihedral);
    e->dihedrals[e->nr_dihedrals].i = i;
    e->dihedrals[e->nr_dihedrals].j = j;
    e->dihedrals[e->nr_dihedrals].k = k;
    e->dihedrals[e->nr_dihedrals].l = l;
    e->dihedrals[e->nr_dihedrals].pid = pid;
    e->nr_dihedrals += 1;
    return engine_err_ok;
}

int engine_gettype(struct engine *e, char *name) {
    if (e == NULL || name == NULL) return eerror_(engine_err_null);
    for (int k = 0; k < e->nr_types; k++)
        if (strcmp(e->types[k].name, name) == 0) return k;
    return engine_err_range;
}

int engine_gettype2(struct engine *e, char *name2) {
    if (e == NULL || name2 == NULL) return eerror_(engine_err_null);
    for (int k = 0; k < e->nr_types; k++)
        if (strcmp(e->types[k].name2, name2) == 0) return k;
    return engine_err_range;
}

struct potential *engine_getpot(struct engine *e, int i, int j) {
    if (e == NULL || i < 0 || j < 0 || i >= e->max_type || j >= e->max_type) return NULL;
    return e->p[i * e->max_type + j];
}

int engine_gettypeinfo(struct engine *e, int k, double *out, char *name, char *name2) {
    if (e == NULL) return eerror_(engine_err_null);
    if (k < 0 || k >= e->nr_types) return eerror_(engine_err_range);
    if (out) {
        out[0] = e->types[k].mass;
        out[1] = e->types[k].charge;
        out[2] = e->types[k].eps;
        out[3] = e->types[k].rmin;
    }
    if (name) snprintf(name, 64, "%s", e->types[k].name);
    if (name2) snprintf(name2, 64, "%s", e->types[k].name2);
    return e->nr_types;
}

int engine_gettopology(struct engine *e, int which, int *out) {
    int k;
    if (e == NULL) return eerror_(engine_err_null);
    if (which == 0) {
        if (out)
            for (k = 0; k < e->nr_bonds; k++) {
                out[2 * k] = e->bonds[k].i;
                out[2 * k + 1] = e->bonds[k].j;
            }
        return e->nr_bonds;
    }
    if (which == 1) {
        if (out)
            for (k = 0; k < e->nr_angles; k++) {
                out[3 * k] = e->angles[k].i;
                out[3 * k + 1] = e->angles[k].j;
                out[3 * k + 2] = e->angles[k].k;
            }
        return e->nr_angles;
    }
    if (which == 2) {
        if (out)
            for (k = 0; k < e->nr_dihedrals; k++) {
                out[5 * k] = e->dihedrals[k].i;
                out[5 * k + 1] = e->dihedrals[k].j;
                out[5 * k + 2] = e->dihedrals[k].k;
                out[5 * k + 3] = e->dihedrals[k].l;
                out[5 * k + 4] = e->dihedrals[k].pid;
            }
        return e->nr_dihedrals;
    }
    return eerror_(engine_err_range);
}

/* ---------------------------------------------------------------- helpers */

#define TOK 100

static int get_double(struct reader *r, double *v) {
    char buff[TOK], *end;
    if (reader_gettoken(r, buff, TOK) < 0) return engine_err_reader;
    *v = strtod(buff, &end);
    return (*end != 0) ? engine_err_cpf : engine_err_ok;
}

static int get_long(struct reader *r, long *v, int err) {
    char buff[TOK], *end;
    if (reader_gettoken(r, buff, TOK) < 0) return engine_err_reader;
    *v = strtol(buff, &end, 0);
    return (*end != 0) ? err : engine_err_ok;
}

/* Read n tokens (types of a bonded term) and then m numbers. */
static int skip_term(struct reader *r, char types[][TOK], int n, double *par, int m) {
    int k, rc;
    for (k = 0; k < n; k++)
        if (reader_gettoken(r, types[k], TOK) < 0) return engine_err_reader;
    for (k = 0; k < m; k++)
        if ((rc = get_double(r, &par[k])) < 0) return rc;
    return engine_err_ok;
}

/* Bonds between particles of types t1 and t2 (either order) become rigid
   constraints of length d and leave the bond list (engine_io.c:696-716). */
static int rigid_bonds(struct engine *e, const char *t1, const char *t2, double d) {
    for (int k = 0; k < e->nr_bonds; k++) {
        struct part *pi = e->s.partlist[e->bonds[k].i], *pj = e->s.partlist[e->bonds[k].j];
        const char *ni, *nj;
        if (pi == NULL || pj == NULL) continue;
        ni = e->types[pi->type].name;
        nj = e->types[pj->type].name;
        if ((strcmp(ni, t1) == 0 && strcmp(nj, t2) == 0) || (strcmp(ni, t2) == 0 && strcmp(nj, t1) == 0)) {
            if (engine_rigid_add(e, e->bonds[k].i, e->bonds[k].j, d) < 0) return eerror_(engine_err);
            e->nr_bonds -= 1;
            e->bonds[k] = e->bonds[e->nr_bonds];
            k -= 1;
        }
    }
    return engine_err_ok;
}

/* Inner radius: bisection on [r0/2, r0] for |V| = engine_maxKcutoff |V(r0)|. */
static double inner_radius(double r0, double A, double B, double q, double kappa) {
    double al = r0 / 2, ar = r0, am, vm, Kr0;
    if (kappa < 0.0)
        Kr0 = fabs(potential_LJ126(r0, A, B));
    else if (kappa == 0.0)
        Kr0 = fabs(potential_LJ126(r0, A, B) + potential_escale * q / r0);
    else
        Kr0 = fabs(potential_LJ126(r0, A, B) + q * potential_Ewald(r0, kappa));
    while (ar - al > 1e-5) {
        am = 0.5 * (al + ar);
        if (kappa < 0.0)
            vm = potential_LJ126(am, A, B);
        else if (kappa == 0.0)
            vm = potential_LJ126(am, A, B) + potential_escale * q / am;
        else
            vm = potential_LJ126(am, A, B) + q * potential_Ewald(am, kappa);
        if (fabs(vm) < engine_maxKcutoff * Kr0)
            ar = am;
        else
            al = am;
    }
    return al;
}

/* The nonbonded tables of every type pair (engine_io.c:526-612 for xplor,
   1106-1220 for cpf; switched: the cpf forms). */
static int nb_tables(struct engine *e, double kappa, double tol, int switched, double switchlen) {
    const double rc = e->s.cutoff;
    for (int j = 0; j < e->nr_types; j++)
        for (int k = j; k < e->nr_types; k++) {
            const struct part_type *tj = &e->types[j], *tk = &e->types[k];
            struct potential *p = NULL;
            int jj, kk;
            if ((tj->eps == 0.0 || tk->eps == 0.0) && (kappa < 0.0 || tj->charge == 0.0 || tk->charge == 0.0))
                continue;
            /* an earlier pair of types with the same parameters */
            for (jj = 0; jj < j; jj++) {
                const struct part_type *t = &e->types[jj];
                if (t->eps == tj->eps && t->rmin == tj->rmin && (kappa < 0.0 || t->charge == tj->charge)) break;
            }
            for (kk = 0; kk < k; kk++) {
                const struct part_type *t = &e->types[kk];
                if (t->eps == tk->eps && t->rmin == tk->rmin && (kappa < 0.0 || t->charge == tk->charge)) break;
            }
            if (jj < j && kk < k) {
                if (e->p[jj + e->max_type * kk] != NULL && engine_addpot(e, e->p[jj + e->max_type * kk], j, k) < 0)
                    return eerror_(engine_err);
                continue;
            }
            const double K = sqrt(tj->eps * tk->eps), r0 = tj->rmin + tk->rmin;
            const double r2 = r0 * r0, r6 = r2 * r2 * r2;
            const double A = K * r6 * r6, B = 2 * K * r6, q = tj->charge * tk->charge;
            /* As in the reference, a pair without a Lennard-Jones well first
               gets a pure electrostatic table, which the general table below
               then replaces. */
            if (K == 0.0 && q != 0.0 && kappa >= 0.0) {
                p = (kappa > 0.0) ? potential_create_Ewald(0.1, rc, q, kappa, tol) : potential_create_Coulomb(0.1, rc, q, tol);
                if (p == NULL) return eerror_(engine_err_potential);
                potential_free(p);
                p = NULL;
            }
            const double al = inner_radius(r0, A, B, q, kappa);
            if (kappa < 0.0)
                p = switched ? potential_create_LJ126_switch(al, rc, A, B, switchlen, tol)
                             : potential_create_LJ126(al, rc, A, B, tol);
            else if (kappa == 0.0)
                p = potential_create_LJ126_Coulomb(al, rc, A, B, q, tol);
            else
                p = switched ? potential_create_LJ126_Ewald_switch(al, rc, A, B, q, kappa, switchlen, tol)
                             : potential_create_LJ126_Ewald(al, rc, A, B, q, kappa, tol);
            if (p == NULL) return eerror_(engine_err_potential);
            if (engine_addpot(e, p, j, k) < 0) return eerror_(engine_err);
        }
    return engine_err_ok;
}

/* Store eps (kcal/mol -> kJ/mol) and rmin (rscale: A -> nm) for every type
   called name. */
static void set_lj(struct engine *e, const char *name, double K, double r0, double rscale) {
    for (int k = 0; k < e->nr_types; k++)
        if (strcmp(e->types[k].name, name) == 0) {
            e->types[k].eps = 4.184 * K;
            e->types[k].rmin = rscale * r0;
        }
}

/* ---------------------------------------------------------------- readers */

int engine_read_cpf(struct engine *e, int cpf, double kappa, double tol, int rigidH, double switchlen) {
    struct reader r;
    char buff[TOK], t[4][TOK];
    double par[3];
    int rc = engine_err_ok;
    if (e == NULL) return eerror_(engine_err_null);
    if (reader_init(&r, cpf, NULL, "!", "\n", engine_readbuff) < 0) return eerror_(engine_err_reader);
    /* title lines start with '*' */
    while (r.c == '*')
        if (reader_skipline(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
    if (reader_gettoken(&r, buff, TOK) < 0) {
        rc = engine_err_reader;
        goto fail;
    }
    if (strncmp(buff, "BOND", 4) != 0) {
        rc = engine_err_cpf;
        goto fail;
    }
    /* bonds: type1 type2 K r0, up to ANGL */
    for (;;) {
        if (reader_gettoken(&r, t[0], TOK) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        if (strncmp(t[0], "ANGL", 4) == 0) break;
        if (reader_gettoken(&r, t[1], TOK) < 0 || get_double(&r, &par[0]) < 0 || get_double(&r, &par[1]) < 0) {
            rc = engine_err_cpf;
            goto fail;
        }
        if (rigidH && (t[0][0] == 'H' || t[1][0] == 'H') && (rc = rigid_bonds(e, t[0], t[1], 0.1 * par[1])) < 0)
            goto fail;
        if (reader_skipline(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
    }
    /* angles (three types, K, theta0, optional Urey-Bradley), dihedrals (four
       types, K, n, delta), impropers (four types, K, 0, psi0): read past */
    {
        static const char *next[3] = {"DIHE", "IMPR", "NONB"};
        static const int ntype[3] = {3, 4, 4}, npar[3] = {2, 3, 3};
        for (int sec = 0; sec < 3; sec++) {
            if (reader_skipline(&r) < 0) {
                rc = engine_err_reader;
                goto fail;
            }
            for (;;) {
                if (reader_gettoken(&r, t[0], TOK) < 0) {
                    rc = engine_err_reader;
                    goto fail;
                }
                if (strncmp(t[0], next[sec], 4) == 0) break;
                if ((rc = skip_term(&r, &t[1], ntype[sec] - 1, par, npar[sec])) < 0) goto fail;
                if (reader_skipline(&r) < 0) {
                    rc = engine_err_reader;
                    goto fail;
                }
            }
        }
    }
    /* nonbonded: type ignored eps Rmin/2 [eps,1-4 Rmin/2,1-4], up to HBOND */
    if (reader_skipline(&r) < 0) {
        rc = engine_err_reader;
        goto fail;
    }
    for (;;) {
        if (reader_gettoken(&r, t[0], TOK) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        if (strncmp(t[0], "HBOND", 5) == 0) break;
        if (reader_skiptoken(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        if (get_double(&r, &par[0]) < 0 || get_double(&r, &par[1]) < 0) {
            rc = engine_err_cpf;
            goto fail;
        }
        set_lj(e, t[0], par[0], par[1], 0.1);
        if (reader_skipline(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
    }
    reader_close(&r);
    return nb_tables(e, kappa, tol, 1, switchlen);
fail:
    reader_close(&r);
    return eerror_(rc);
}

int engine_read_xplor(struct engine *e, int xplor, double kappa, double tol, int rigidH) {
    struct reader r;
    char buff[TOK], t[4][TOK];
    double par[3];
    int res, rc = engine_err_ok;
    if (e == NULL) return eerror_(engine_err_null);
    if (reader_init(&r, xplor, NULL, "!{", "\n", engine_readbuff) < 0) return eerror_(engine_err_reader);
    while (!(r.flags & reader_flag_eof)) {
        if ((res = reader_gettoken(&r, buff, TOK)) == reader_err_eof) break;
        if (res < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        if (strncasecmp(buff, "BOND", 4) == 0) {
            if ((rc = skip_term(&r, t, 2, par, 2)) < 0) goto fail;
            if (rigidH && (t[0][0] == 'H' || t[1][0] == 'H') && (rc = rigid_bonds(e, t[0], t[1], 0.1 * par[1])) < 0)
                goto fail;
        } else if (strncasecmp(buff, "NONB", 4) == 0) {
            /* type eps sigma-like r0 (the reference reads Rmin as 2 x 0.05 r0) */
            if (reader_gettoken(&r, t[0], TOK) < 0 || get_double(&r, &par[0]) < 0 || get_double(&r, &par[1]) < 0) {
                rc = engine_err_cpf;
                goto fail;
            }
            set_lj(e, t[0], par[0], par[1], 0.05);
        }
        /* ANGLe, DIHEdral, IMPRoper and other lines: bonded terms are not
           evaluated here, so the rest of the line is skipped unparsed (the
           reference also parses them, and fails on "MULTIPLE=" dihedrals) */
        if (reader_skipline(&r) < 0) break; /* the last line may end the file */
    }
    reader_close(&r);
    return nb_tables(e, kappa, tol, 0, 0.0);
fail:
    reader_close(&r);
    return eerror_(rc);
}

/* Read a count line "n !NAME" of a PSF section. */
static int psf_count(struct reader *r, const char *name, long *n) {
    char buff[TOK];
    int rc;
    if ((rc = get_long(r, n, engine_err_psf)) < 0) return rc;
    if (reader_getcomment(r, buff, TOK) < 0) return engine_err_reader;
    if (strncmp(buff, name, strlen(name)) != 0) return engine_err_psf;
    return engine_err_ok;
}

/* Read nid particle ids (1-based in the file, 0-based out). */
static int psf_ids(struct reader *r, long *ids, int nid) {
    int rc;
    for (int q = 0; q < nid; q++) {
        if ((rc = get_long(r, &ids[q], engine_err_psf)) < 0) return rc;
        ids[q] -= 1;
    }
    return engine_err_ok;
}

int engine_read_psf(struct engine *e, int psf, int pdb) {
    struct reader r;
    char buff[TOK], type[TOK], name2[2 * TOK], *end;
    int *typeids = NULL, *resids = NULL, rc = engine_err_ok, open = 0;
    long n = 0, k, ids[4];
    if (e == NULL) return eerror_(engine_err_null);
    if (reader_init(&r, psf, NULL, "!", "\n", engine_readbuff) < 0) return eerror_(engine_err_reader);
    open = 1;
    if (reader_gettoken(&r, buff, TOK) < 0 || strcmp(buff, "PSF") != 0) {
        rc = engine_err_psf;
        goto fail;
    }
    /* title: a count, then the rest of its line and that many lines */
    if ((rc = get_long(&r, &n, engine_err_psf)) < 0) goto fail;
    for (k = 0; k <= n; k++)
        if (reader_skipline(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
    if ((rc = psf_count(&r, "NATOM", &n)) < 0) goto fail;
    if ((typeids = (int *)malloc(sizeof(int) * (n > 0 ? n : 1))) == NULL ||
        (resids = (int *)malloc(sizeof(int) * (n > 0 ? n : 1))) == NULL) {
        rc = engine_err_malloc;
        goto fail;
    }
    /* atoms: id segment resid resname atomname type charge mass 0; one type
       per atom type and charge, second name = type + charge text */
    for (k = 0; k < n; k++) {
        long rid;
        double q, m;
        int id;
        if (reader_skiptoken(&r) < 0 || reader_skiptoken(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        if ((rc = get_long(&r, &rid, engine_err_psf)) < 0) goto fail;
        resids[k] = (int)rid;
        if (reader_skiptoken(&r) < 0 || reader_skiptoken(&r) < 0 || reader_gettoken(&r, type, TOK) < 0 ||
            reader_gettoken(&r, buff, TOK) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
        q = strtod(buff, &end);
        if (*end != 0) {
            rc = engine_err_psf;
            goto fail;
        }
        snprintf(name2, sizeof(name2), "%s%s", type, buff);
        if ((rc = get_double(&r, &m)) < 0) {
            rc = engine_err_psf;
            goto fail;
        }
        if ((id = engine_gettype2(e, name2)) >= 0)
            typeids[k] = id;
        else if (id == engine_err_range) {
            if ((typeids[k] = engine_addtype(e, m, q, type, name2)) < 0) {
                rc = engine_err;
                goto fail;
            }
        } else {
            rc = engine_err;
            goto fail;
        }
        if (reader_skiptoken(&r) < 0) {
            rc = engine_err_reader;
            goto fail;
        }
    }
    /* bonds, angles, dihedrals, impropers */
    {
        static const char *names[4] = {"NBOND", "NTHETA", "NPHI", "NIMP"};
        static const int nid[4] = {2, 3, 4, 4};
        for (int sec = 0; sec < 4; sec++) {
            long m;
            if ((rc = psf_count(&r, names[sec], &m)) < 0) goto fail;
            for (long q = 0; q < m; q++) {
                if ((rc = psf_ids(&r, ids, nid[sec])) < 0) goto fail;
                if (sec == 0)
                    rc = engine_bond_add(e, (int)ids[0], (int)ids[1]);
                else if (sec == 1)
                    rc = engine_angle_add(e, (int)ids[0], (int)ids[1], (int)ids[2], -1);
                else
                    rc = engine_dihedral_add(e, (int)ids[0], (int)ids[1], (int)ids[2], (int)ids[3], sec == 2 ? -1 : -2);
                if (rc < 0) {
                    rc = engine_err;
                    goto fail;
                }
            }
        }
    }
    reader_close(&r);
    open = 0;
    /* PDB: REMARK lines, ATOM records (serial, name, two skipped fields,
       x y z in Angstrom), END */
    if (reader_init(&r, pdb, NULL, NULL, NULL, engine_readbuff) < 0) {
        rc = engine_err_reader;
        goto fail;
    }
    open = 1;
    {
        struct part p;
        long pid = 0;
        double x[3];
        memset(&p, 0, sizeof(p));
        for (;;) {
            if (reader_gettoken(&r, buff, TOK) < 0) {
                rc = engine_err_reader;
                goto fail;
            }
            if (strncmp(buff, "REMARK", 6) == 0) {
                if (reader_skipline(&r) < 0) {
                    rc = engine_err_reader;
                    goto fail;
                }
            } else if (strncmp(buff, "ATOM", 4) == 0) {
                if (reader_skiptoken(&r) < 0 || reader_gettoken(&r, type, TOK) < 0 || reader_skiptoken(&r) < 0 ||
                    reader_skiptoken(&r) < 0) {
                    rc = engine_err_reader;
                    goto fail;
                }
                pid += 1;
                if (pid > n) {
                    rc = engine_err_pdb;
                    goto fail;
                }
                for (int d = 0; d < 3; d++) {
                    if (reader_gettoken(&r, buff, TOK) < 0) {
                        rc = engine_err_reader;
                        goto fail;
                    }
                    x[d] = fmod(e->s.dim[d] - e->s.origin[d] + 0.1 * strtod(buff, &end), e->s.dim[d]) + e->s.origin[d];
                    if (*end != 0) {
                        rc = engine_err_pdb;
                        goto fail;
                    }
                }
                p.id = (int)(pid - 1);
                p.vid = resids[pid - 1];
                p.q = (float)e->types[typeids[pid - 1]].charge;
                p.flags = part_flag_none;
                p.type = (short)typeids[pid - 1];
                if (space_addpart(&e->s, &p, x) < 0) {
                    rc = engine_err_space;
                    goto fail;
                }
                if (reader_skipline(&r) < 0) {
                    rc = engine_err_reader;
                    goto fail;
                }
            } else if (strncmp(buff, "END", 3) == 0)
                break;
            else {
                rc = engine_err_pdb;
                goto fail;
            }
        }
    }
    reader_close(&r);
    free(typeids);
    free(resids);
    return engine_err_ok;
fail:
    if (open) reader_close(&r);
    free(typeids);
    free(resids);
    return eerror_(rc);
}
