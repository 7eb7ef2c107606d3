/*
 * TEST ORACLE ONLY — internal interface of the frictional-boundary oracle
 * (diffphd_contact.c) used by oracle_forward in diffphd_oracle.c.
 */
#ifndef CONTACT_ORACLE_H
#define CONTACT_ORACLE_H

#include "diffphd_oracle.h"

typedef struct {
  int K, active;
  double* panel;     /* [K, n] A^{-1} e_node */
  double* block;     /* [K, K] the panel at the registered vertices */
  double* delassus;  /* [K, K] */
  double* reg;       /* [K] */
  double *t1, *t2;   /* [K, 3] */
  double* fl;        /* [K, 2] tangential multipliers */
  double *gap, *diag, *resid, *rhs;
  double *force, *response;  /* [K, 3] */
  int *order, *slot;
  double *matrix, *compact;
  double* surface;   /* [n, 3] the iterate a mesh boundary last saw */
} contact_work;

int contact_prepare(const oracle_problem* p, contact_work* w);
void contact_free(contact_work* w);
void contact_begin(const oracle_problem* p, contact_work* w, const double* pos, double dt);
void contact_iterate(const oracle_problem* p, contact_work* w, const double* x, const double* solved,
                     const double* pos, double* out);
void contact_finish(const oracle_problem* p, contact_work* w, const double* pos, const double* vel, double dt,
                    double* x);
void oracle_ldlt_solve(double* m, int n, double* b);

#endif
