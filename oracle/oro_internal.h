/*
 * oro_internal.h -- solver layout shared by the oracle's translation units
 * (vt_oracle.c, vt_oracle_adjoint.c).  TEST INFRASTRUCTURE ONLY.
 */
#ifndef ORO_INTERNAL_H
#define ORO_INTERNAL_H

#include "vt_oracle.h"

#define MAXNP 10
#define MAXNQ 10

struct oro_solver {
    int ne, p, np, nq;
    double xw[MAXNQ][2];
    double *vert;        /* [ne][6] */
    int *bnd, *type;     /* [ne][3], [ne] */
    double *nodes_phys;  /* [ne][np][2] */
    double *geo;         /* [ne][16]: jac4 detJ invJ4 detInvJ n6 ... */
    double *len;         /* [ne][3] */
    double Minv_ref[MAXNP * MAXNP];
    double S_ref[2 * MAXNP * MAXNP];
    double *Minv_phys;   /* [ne][np][np] */
    double *S_phys;      /* [ne][2][np][np] */
    double plus[3][MAXNP][MAXNQ], minus[3][MAXNP][MAXNQ];
    double *U, *F, *time, *dU;
    int order;
    int err[6];        /* elem, q, side, code, step, stage of the first failure */
    int cur_step, cur_stage;
};

#define GEO_JAC 0
#define GEO_DET 4
#define GEO_INVJ 5
#define GEO_DETINV 9
#define GEO_N 10
#define GEO_STRIDE 16

#endif
