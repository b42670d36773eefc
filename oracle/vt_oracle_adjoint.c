/*
 * vt_oracle_adjoint.c -- derivatives of the oracle's discrete DG + RK4 map:
 * tangents (forward mode) and adjoints (reverse mode), written independently
 * of each other so that the dot-product identity <w, J v> = <J^T w, v> pins
 * one against the other, while central finite differences of the oracle's
 * own forward functions pin the tangents.
 *
 * TEST INFRASTRUCTURE ONLY (the CPU reference gradient); the product path
 * never links or calls this file.
 *
 * The reference has no derivatives; the functions differentiated are the
 * oracle restatements of Numericalflux.cpp:22-166 (Roe flux), Evolve.cpp:135-307
 * (right-hand side, through rhs_element in vt_oracle.c), Timestepping.cpp:85-305
 * (RK4) and Element.cpp:134-166 (vortex initial state).  Branches (entropy fix,
 * |.|) are differentiated on the side the forward takes.
 */
#define _DEFAULT_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oro_internal.h"
#include "vt_oracle.h"

/* ---- dual numbers for the tangents --------------------------------------- */

typedef struct { double v, d; } dual;

static inline dual D(double v, double d) { dual r = {v, d}; return r; }
static inline dual dc(double v) { return D(v, 0.0); }
static inline dual dadd(dual a, dual b) { return D(a.v + b.v, a.d + b.d); }
static inline dual dsub(dual a, dual b) { return D(a.v - b.v, a.d - b.d); }
static inline dual dmul(dual a, dual b) { return D(a.v * b.v, a.d * b.v + a.v * b.d); }
static inline dual dscale(double s, dual a) { return D(s * a.v, s * a.d); }
static inline dual ddiv(dual a, dual b) { return D(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
static inline dual dsqrt(dual a) { double r = sqrt(a.v); return D(r, 0.5 * a.d / r); }
static inline dual dabs(dual a) { return a.v < 0 ? D(-a.v, -a.d) : a; }

/* Roe flux of Numericalflux.cpp:31-156 on dual numbers. */
static void roe_dual(const dual *UL, const dual *UR, const double *nv, dual *F) {
    const double gmi = 0.4;
    dual nx = dc(nv[0]), ny = dc(nv[1]);
    dual rL = UL[0], uL = ddiv(UL[1], rL), vL = ddiv(UL[2], rL);
    dual unL = dadd(dmul(uL, nx), dmul(vL, ny));
    dual pL = dscale(gmi, dsub(UL[3], dscale(0.5, ddiv(dadd(dmul(UL[1], UL[1]), dmul(UL[2], UL[2])), rL))));
    dual rHL = dadd(UL[3], pL), HL = ddiv(rHL, rL);
    dual rR = UR[0], uR = ddiv(UR[1], rR), vR = ddiv(UR[2], rR);
    dual unR = dadd(dmul(uR, nx), dmul(vR, ny));
    dual pR = dscale(gmi, dsub(UR[3], dscale(0.5, ddiv(dadd(dmul(UR[1], UR[1]), dmul(UR[2], UR[2])), rR))));
    dual rHR = dadd(UR[3], pR), HR = ddiv(rHR, rR);
    dual FL[4] = {dmul(rL, unL), dadd(dmul(UL[1], unL), dmul(pL, nx)), dadd(dmul(UL[2], unL), dmul(pL, ny)),
                  dmul(rHL, unL)};
    dual FR[4] = {dmul(rR, unR), dadd(dmul(UR[1], unR), dmul(pR, nx)), dadd(dmul(UR[2], unR), dmul(pR, ny)),
                  dmul(rHR, unR)};
    dual du[4];
    for (int k = 0; k < 4; ++k) du[k] = dsub(UR[k], UL[k]);
    dual di = dsqrt(ddiv(rR, rL));
    dual d1 = ddiv(dc(1.0), dadd(dc(1.0), di));
    dual ui = dmul(dadd(dmul(di, uR), uL), d1);
    dual vi = dmul(dadd(dmul(di, vR), vL), d1);
    dual Hi = dmul(dadd(dmul(di, HR), HL), d1);
    dual af = dscale(0.5, dadd(dmul(ui, ui), dmul(vi, vi)));
    dual ucp = dadd(dmul(ui, nx), dmul(vi, ny));
    dual c2 = dscale(gmi, dsub(Hi, af));
    dual ci = dmul(c2, c2);
    dual ci1 = ddiv(dc(1.0), ci);
    dual eps = dscale(0.1, ci);
    dual l[3] = {dadd(ucp, ci), dsub(ucp, ci), ucp};
    for (int i = 0; i < 3; ++i) {
        if (l[i].v < eps.v && l[i].v > -eps.v) l[i] = dscale(0.5, dadd(eps, ddiv(dmul(l[i], l[i]), eps)));
        l[i] = dabs(l[i]);
    }
    dual l3 = l[2];
    dual s1 = dscale(0.5, dadd(l[0], l[1])), s2 = dscale(0.5, dsub(l[0], l[1]));
    dual G1 = dscale(gmi, dadd(dsub(dsub(dmul(af, du[0]), dmul(ui, du[1])), dmul(vi, du[2])), du[3]));
    dual G2 = dadd(dadd(dscale(-1.0, dmul(ucp, du[0])), dmul(du[1], nx)), dmul(du[2], ny));
    dual C1 = dadd(dmul(dmul(dmul(G1, dsub(s1, l3)), ci1), ci1), dmul(dmul(G2, s2), ci1));
    dual C2 = dadd(dmul(dmul(G1, s2), ci1), dmul(G2, dsub(s1, l3)));
    dual one = dc(1.0);
    dual a[4] = {one, ui, vi, Hi}, b[4] = {dc(0.0), nx, ny, ucp};
    for (int k = 0; k < 4; ++k)
        F[k] = dsub(dscale(0.5, dadd(FL[k], FR[k])),
                    dscale(0.5, dadd(dadd(dmul(l3, du[k]), dmul(C1, a[k])), dmul(C2, b[k]))));
}

void oro_roe_flux_jvp(const double *UL, const double *UR, const double *n, const double *dUL, const double *dUR,
                      double *dF) {
    dual L[4], R[4], F[4];
    for (int k = 0; k < 4; ++k) { L[k] = D(UL[k], dUL[k]); R[k] = D(UR[k], dUR[k]); }
    roe_dual(L, R, n, F);
    for (int k = 0; k < 4; ++k) dF[k] = F[k].d;
}

/* ---- Roe flux adjoint (reverse sweep written out by hand) ----------------- */

/* Per side: primitive quantities and their adjoint accumulation. */
typedef struct { double r, u, v, un, p, rH, H; } side_t;

static side_t side_values(const double *U, double nx, double ny) {
    side_t s;
    s.r = U[0];
    s.u = U[1] / s.r;
    s.v = U[2] / s.r;
    s.un = s.u * nx + s.v * ny;
    s.p = 0.4 * (U[3] - 0.5 * (U[1] * U[1] + U[2] * U[2]) / s.r);
    s.rH = U[3] + s.p;
    s.H = s.rH / s.r;
    return s;
}

/* Adjoint of one side: given the adjoints of F_side (bFs[4]), u, v, H and the
 * density r, accumulate into bU. */
static void side_adjoint(const double *U, const side_t *s, double nx, double ny, const double *bFs, double bu,
                         double bv, double bH, double br, double *bU) {
    const double gmi = 0.4;
    /* F = (r un, U1 un + p nx, U2 un + p ny, rH un) */
    br += bFs[0] * s->un;
    double bun = bFs[0] * s->r + bFs[1] * U[1] + bFs[2] * U[2] + bFs[3] * s->rH;
    double bU1 = bFs[1] * s->un, bU2 = bFs[2] * s->un;
    double bp = bFs[1] * nx + bFs[2] * ny;
    double brH = bFs[3] * s->un;
    /* H = rH / r */
    brH += bH / s->r;
    br -= bH * s->H / s->r;
    /* rH = U3 + p */
    double bU3 = brH;
    bp += brH;
    /* un = u nx + v ny */
    bu += bun * nx;
    bv += bun * ny;
    /* p = gmi (U3 - 0.5 (U1^2 + U2^2) / r) */
    bU3 += gmi * bp;
    bU1 -= gmi * U[1] / s->r * bp;
    bU2 -= gmi * U[2] / s->r * bp;
    br += gmi * 0.5 * (U[1] * U[1] + U[2] * U[2]) / (s->r * s->r) * bp;
    /* u = U1 / r, v = U2 / r */
    bU1 += bu / s->r;
    br -= bu * s->u / s->r;
    bU2 += bv / s->r;
    br -= bv * s->v / s->r;
    bU[0] += br;
    bU[1] += bU1;
    bU[2] += bU2;
    bU[3] += bU3;
}

void oro_roe_flux_vjp(const double *UL, const double *UR, const double *n, const double *bF, double *bUL,
                      double *bUR) {
    const double gmi = 0.4, nx = n[0], ny = n[1];
    /* forward sweep */
    side_t L = side_values(UL, nx, ny), R = side_values(UR, nx, ny);
    double du[4];
    for (int k = 0; k < 4; ++k) du[k] = UR[k] - UL[k];
    double di = sqrt(R.r / L.r);
    double d1 = 1.0 / (1.0 + di);
    double ui = (di * R.u + L.u) * d1, vi = (di * R.v + L.v) * d1, Hi = (di * R.H + L.H) * d1;
    double af = 0.5 * (ui * ui + vi * vi);
    double ucp = ui * nx + vi * ny;
    double c2 = gmi * (Hi - af);
    double ci = c2 * c2, ci1 = 1.0 / ci, eps = 0.1 * ci;
    double lr[3] = {ucp + ci, ucp - ci, ucp}, lf[3], la[3];
    int fixed[3];
    for (int i = 0; i < 3; ++i) {
        fixed[i] = (lr[i] < eps) && (lr[i] > -eps);
        lf[i] = fixed[i] ? 0.5 * (eps + lr[i] * lr[i] / eps) : lr[i];
        la[i] = fabs(lf[i]);
    }
    double l3 = la[2], s1 = 0.5 * (la[0] + la[1]), s2 = 0.5 * (la[0] - la[1]);
    double G1 = gmi * (af * du[0] - ui * du[1] - vi * du[2] + du[3]);
    double G2 = -ucp * du[0] + du[1] * nx + du[2] * ny;
    double C1 = G1 * (s1 - l3) * ci1 * ci1 + G2 * s2 * ci1;
    double C2 = G1 * s2 * ci1 + G2 * (s1 - l3);
    /* reverse sweep */
    double bFL[4], bFR[4], bdu[4];
    double bl3 = 0.0, bC1 = 0.0, bC2 = 0.0, bui = 0.0, bvi = 0.0, bHi = 0.0, bucp = 0.0;
    for (int k = 0; k < 4; ++k) {
        bFL[k] = 0.5 * bF[k];
        bFR[k] = 0.5 * bF[k];
        bl3 -= 0.5 * bF[k] * du[k];
        bdu[k] = -0.5 * l3 * bF[k];
    }
    bC1 = -0.5 * (bF[0] + bF[1] * ui + bF[2] * vi + bF[3] * Hi);
    bC2 = -0.5 * (bF[1] * nx + bF[2] * ny + bF[3] * ucp);
    bui -= 0.5 * C1 * bF[1];
    bvi -= 0.5 * C1 * bF[2];
    bHi -= 0.5 * C1 * bF[3];
    bucp -= 0.5 * C2 * bF[3];
    double bG1 = 0.0, bG2 = 0.0, bs1 = 0.0, bs2 = 0.0, bci1 = 0.0;
    /* C2 = G1 s2 ci1 + G2 (s1 - l3) */
    bG1 += bC2 * s2 * ci1;
    bs2 += bC2 * G1 * ci1;
    bci1 += bC2 * G1 * s2;
    bG2 += bC2 * (s1 - l3);
    bs1 += bC2 * G2;
    bl3 -= bC2 * G2;
    /* C1 = G1 (s1 - l3) ci1^2 + G2 s2 ci1 */
    bG1 += bC1 * (s1 - l3) * ci1 * ci1;
    bs1 += bC1 * G1 * ci1 * ci1;
    bl3 -= bC1 * G1 * ci1 * ci1;
    bci1 += bC1 * (2.0 * G1 * (s1 - l3) * ci1 + G2 * s2);
    bG2 += bC1 * s2 * ci1;
    bs2 += bC1 * G2 * ci1;
    /* G2 = -ucp du0 + du1 nx + du2 ny */
    bucp -= bG2 * du[0];
    bdu[0] -= bG2 * ucp;
    bdu[1] += bG2 * nx;
    bdu[2] += bG2 * ny;
    /* G1 = gmi (af du0 - ui du1 - vi du2 + du3) */
    double baf = bG1 * gmi * du[0];
    bdu[0] += bG1 * gmi * af;
    bui -= bG1 * gmi * du[1];
    bdu[1] -= bG1 * gmi * ui;
    bvi -= bG1 * gmi * du[2];
    bdu[2] -= bG1 * gmi * vi;
    bdu[3] += bG1 * gmi;
    /* s1, s2, l3 from |l| */
    double bla[3] = {0.5 * bs1 + 0.5 * bs2, 0.5 * bs1 - 0.5 * bs2, bl3};
    double bl[3], beps = 0.0, bci = 0.0;
    for (int i = 0; i < 3; ++i) {
        double blf = lf[i] < 0 ? -bla[i] : bla[i];
        if (fixed[i]) {
            bl[i] = blf * lr[i] / eps;
            beps += blf * (0.5 - 0.5 * lr[i] * lr[i] / (eps * eps));
        } else {
            bl[i] = blf;
        }
    }
    /* l = (ucp + ci, ucp - ci, ucp) */
    bucp += bl[0] + bl[1] + bl[2];
    bci += bl[0] - bl[1];
    /* eps = 0.1 ci, ci1 = 1/ci, ci = c2^2 */
    bci += 0.1 * beps;
    bci -= bci1 * ci1 * ci1;
    double bc2 = 2.0 * c2 * bci;
    /* c2 = gmi (Hi - af) */
    bHi += gmi * bc2;
    baf -= gmi * bc2;
    /* ucp = ui nx + vi ny, af = 0.5 (ui^2 + vi^2) */
    bui += bucp * nx + baf * ui;
    bvi += bucp * ny + baf * vi;
    /* Roe averages */
    double bdi = 0.0, bd1 = 0.0;
    double buR = bui * di * d1, buL = bui * d1, bvR = bvi * di * d1, bvL = bvi * d1;
    double bHR = bHi * di * d1, bHL = bHi * d1;
    bdi += (bui * R.u + bvi * R.v + bHi * R.H) * d1;
    bd1 += bui * (di * R.u + L.u) + bvi * (di * R.v + L.v) + bHi * (di * R.H + L.H);
    bdi -= bd1 * d1 * d1;
    /* di = sqrt(rR / rL) */
    double brR = bdi * 0.5 / (di * L.r);
    double brL = -bdi * 0.5 * di / L.r;
    for (int k = 0; k < 4; ++k) { bUL[k] = -bdu[k]; bUR[k] = bdu[k]; }
    side_adjoint(UL, &L, nx, ny, bFL, buL, bvL, bHL, brL, bUL);
    side_adjoint(UR, &R, nx, ny, bFR, buR, bvR, bHR, brR, bUR);
}

/* ---- Euler flux (nodal flux, Timestepping.cpp:59-76) --------------------- */

static void euler_flux_jvp(const double *U, const double *dU, double *dFx, double *dFy) {
    dual u[4], F[8];
    for (int k = 0; k < 4; ++k) u[k] = D(U[k], dU[k]);
    dual rho = u[0], vx = ddiv(u[1], rho), vy = ddiv(u[2], rho), E = ddiv(u[3], rho);
    dual p = dmul(dscale(0.4, rho), dsub(E, dscale(0.5, dadd(dmul(vx, vx), dmul(vy, vy)))));
    dual H = dadd(E, ddiv(p, rho));
    F[0] = dmul(rho, vx);
    F[1] = dadd(dmul(rho, dmul(vx, vx)), p);
    F[2] = dmul(rho, dmul(vx, vy));
    F[3] = dmul(rho, dmul(vx, H));
    F[4] = dmul(rho, vy);
    F[5] = dmul(rho, dmul(vx, vy));
    F[6] = dadd(dmul(rho, dmul(vy, vy)), p);
    F[7] = dmul(rho, dmul(vy, H));
    for (int k = 0; k < 4; ++k) { dFx[k] = F[k].d; dFy[k] = F[4 + k].d; }
}

/* Adjoint of the conservative form Fx = (m1, m1 u + p, m1 v, (E + p) u),
 * Fy = (m2, m2 u, m2 v + p, (E + p) v) with u = m1/rho, v = m2/rho,
 * p = 0.4 (E - 0.5 (m1^2 + m2^2)/rho). */
static void euler_flux_vjp(const double *U, const double *bFx, const double *bFy, double *bU) {
    const double gmi = 0.4;
    double rho = U[0], m1 = U[1], m2 = U[2], E = U[3];
    double u = m1 / rho, v = m2 / rho;
    double p = gmi * (E - 0.5 * (m1 * m1 + m2 * m2) / rho);
    double bm1 = bFx[0] + bFx[1] * u + bFx[2] * v;
    double bm2 = bFy[0] + bFy[1] * u + bFy[2] * v;
    double bu = bFx[1] * m1 + bFx[3] * (E + p) + bFy[1] * m2;
    double bv = bFx[2] * m1 + bFy[2] * m2 + bFy[3] * (E + p);
    double bp = bFx[1] + bFx[3] * u + bFy[2] + bFy[3] * v;
    double bE = bFx[3] * u + bFy[3] * v;
    double brho = 0.0;
    bE += gmi * bp;
    bm1 -= gmi * u * bp;
    bm2 -= gmi * v * bp;
    brho += gmi * 0.5 * (u * u + v * v) * bp;
    bm1 += bu / rho;
    brho -= bu * u / rho;
    bm2 += bv / rho;
    brho -= bv * v / rho;
    bU[0] += brho;
    bU[1] += bm1;
    bU[2] += bm2;
    bU[3] += bE;
}

/* ---- right-hand side f(U) = Minv_phys (V - L) ---------------------------- */

static void side_neighbours(const oro_solver *s, int e, int *nb) {
    nb[0] = s->bnd[3 * e + 2]; /* side 1: vertical (Evolve.cpp:167-202) */
    if (s->type[e] == 0) { nb[1] = s->bnd[3 * e]; nb[2] = s->bnd[3 * e + 1]; }
    else { nb[1] = s->bnd[3 * e + 1]; nb[2] = s->bnd[3 * e]; }
}

void oro_rhs_jvp(const oro_solver *s, const double *dU, double *df) {
    const int np = s->np, nq = s->nq, ne = s->ne;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        int nb[3];
        side_neighbours(s, e, nb);
        const double *Ue = s->U + (size_t)e * np * 4, *dUe = dU + (size_t)e * np * 4;
        const double *nrm = s->geo + e * GEO_STRIDE + GEO_N, *len = s->len + 3 * e;
        double dL[MAXNP][4] = {{0}}, dV[MAXNP][4] = {{0}};
        for (int side = 0; side < 3; ++side)
            for (int q = 0; q < nq; ++q) {
                double Up[4] = {0}, Um[4] = {0}, dUp[4] = {0}, dUm[4] = {0}, dF[4];
                const double *Un = s->U + (size_t)nb[side] * np * 4, *dUn = dU + (size_t)nb[side] * np * 4;
                for (int j = 0; j < np; ++j)
                    for (int k = 0; k < 4; ++k) {
                        Up[k] += s->plus[side][j][q] * Ue[j * 4 + k];
                        dUp[k] += s->plus[side][j][q] * dUe[j * 4 + k];
                        Um[k] += s->minus[side][j][q] * Un[j * 4 + k];
                        dUm[k] += s->minus[side][j][q] * dUn[j * 4 + k];
                    }
                oro_roe_flux_jvp(Up, Um, nrm + 2 * side, dUp, dUm, dF);
                for (int i = 0; i < np; ++i)
                    for (int k = 0; k < 4; ++k) dL[i][k] += len[side] * s->plus[side][i][q] * dF[k] * s->xw[q][1];
            }
        const double *Sx = s->S_phys + (size_t)e * 2 * np * np, *Sy = Sx + np * np;
        for (int j = 0; j < np; ++j) {
            double dFx[4], dFy[4];
            euler_flux_jvp(Ue + 4 * j, dUe + 4 * j, dFx, dFy);
            for (int i = 0; i < np; ++i)
                for (int k = 0; k < 4; ++k) dV[i][k] += Sx[i * np + j] * dFx[k] + Sy[i * np + j] * dFy[k];
        }
        const double *Mi = s->Minv_phys + (size_t)e * np * np;
        for (int i = 0; i < np; ++i)
            for (int k = 0; k < 4; ++k) {
                double a = 0.0;
                for (int j = 0; j < np; ++j) a += Mi[i * np + j] * (dV[j][k] - dL[j][k]);
                df[((size_t)e * np + i) * 4 + k] = a;
            }
    }
}

void oro_rhs_vjp(const oro_solver *s, const double *w, double *g) {
    const int np = s->np, nq = s->nq, ne = s->ne;
    memset(g, 0, sizeof(double) * (size_t)ne * np * 4);
    for (int e = 0; e < ne; ++e) {
        int nb[3];
        side_neighbours(s, e, nb);
        const double *Ue = s->U + (size_t)e * np * 4, *we = w + (size_t)e * np * 4;
        const double *Mi = s->Minv_phys + (size_t)e * np * np;
        const double *nrm = s->geo + e * GEO_STRIDE + GEO_N, *len = s->len + 3 * e;
        double *ge = g + (size_t)e * np * 4;
        /* dU = Mi R:  bR_j = sum_i Mi_ij w_i; R = V - L */
        double bR[MAXNP][4];
        for (int j = 0; j < np; ++j)
            for (int k = 0; k < 4; ++k) {
                double a = 0.0;
                for (int i = 0; i < np; ++i) a += Mi[i * np + j] * we[i * 4 + k];
                bR[j][k] = a;
            }
        /* volume: V_i = sum_j Sx_ij Fx_j + Sy_ij Fy_j */
        const double *Sx = s->S_phys + (size_t)e * 2 * np * np, *Sy = Sx + np * np;
        for (int j = 0; j < np; ++j) {
            double bFx[4] = {0}, bFy[4] = {0};
            for (int i = 0; i < np; ++i)
                for (int k = 0; k < 4; ++k) {
                    bFx[k] += Sx[i * np + j] * bR[i][k];
                    bFy[k] += Sy[i * np + j] * bR[i][k];
                }
            euler_flux_vjp(Ue + 4 * j, bFx, bFy, ge + 4 * j);
        }
        /* surface: L_i = sum_{s,q} len_s plus_s,i(q) F_s,q w_q, with weight -1 in R */
        for (int side = 0; side < 3; ++side) {
            const double *Un = s->U + (size_t)nb[side] * np * 4;
            double *gn = g + (size_t)nb[side] * np * 4;
            for (int q = 0; q < nq; ++q) {
                double Up[4] = {0}, Um[4] = {0}, bF[4] = {0}, bUp[4], bUm[4];
                for (int j = 0; j < np; ++j)
                    for (int k = 0; k < 4; ++k) {
                        Up[k] += s->plus[side][j][q] * Ue[j * 4 + k];
                        Um[k] += s->minus[side][j][q] * Un[j * 4 + k];
                    }
                for (int i = 0; i < np; ++i)
                    for (int k = 0; k < 4; ++k) bF[k] -= len[side] * s->plus[side][i][q] * s->xw[q][1] * bR[i][k];
                oro_roe_flux_vjp(Up, Um, nrm + 2 * side, bF, bUp, bUm);
                for (int j = 0; j < np; ++j)
                    for (int k = 0; k < 4; ++k) {
                        ge[j * 4 + k] += s->plus[side][j][q] * bUp[k];
                        gn[j * 4 + k] += s->minus[side][j][q] * bUm[k];
                    }
            }
        }
    }
}

/* ---- RK4 tangent and adjoint (Timestepping.cpp:85-305) ------------------- */

static void rhs_at(oro_solver *s, const double *U, double *dUdt) {
    oro_set_state(s, U);
    oro_evolve_elem(s, dUdt);
}

int oro_rk4_tangent(oro_solver *s, double dt, int nsteps, const double *dU0, double *dUT) {
    size_t n = (size_t)s->ne * s->np * 4;
    double *U = malloc(sizeof(double) * n), *Us = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n);
    double *k[4], *dk[4], *dv = malloc(sizeof(double) * n), *dUs = malloc(sizeof(double) * n);
    for (int i = 0; i < 4; ++i) { k[i] = malloc(sizeof(double) * n); dk[i] = malloc(sizeof(double) * n); }
    memcpy(U, s->U, sizeof(double) * n);
    memcpy(dv, dU0, sizeof(double) * n);
    const double c[4] = {0.0, 0.5, 0.5, 1.0};
    for (int a = 0; a < nsteps; ++a) {
        for (int st = 0; st < 4; ++st) {
            for (size_t m = 0; m < n; ++m) {
                Us[m] = st ? U[m] + (st == 3 ? k[2][m] : k[st - 1][m] / 2.0) : U[m];
                dUs[m] = st ? dv[m] + c[st] * dk[st - 1][m] : dv[m];
            }
            rhs_at(s, Us, f);
            for (size_t m = 0; m < n; ++m) k[st][m] = dt * f[m];
            oro_rhs_jvp(s, dUs, f);
            for (size_t m = 0; m < n; ++m) dk[st][m] = dt * f[m];
        }
        for (size_t m = 0; m < n; ++m) {
            U[m] = U[m] + (1.0 / 6.0) * (k[0][m] + 2 * k[1][m] + 2 * k[2][m] + k[3][m]);
            dv[m] = dv[m] + (1.0 / 6.0) * (dk[0][m] + 2 * dk[1][m] + 2 * dk[2][m] + dk[3][m]);
        }
    }
    oro_set_state(s, U);
    memcpy(dUT, dv, sizeof(double) * n);
    for (int i = 0; i < 4; ++i) { free(k[i]); free(dk[i]); }
    free(U); free(Us); free(f); free(dv); free(dUs);
    return 0;
}

int oro_rk4_adjoint(oro_solver *s, double dt, int nsteps, const double *lamT, double *lam0) {
    size_t n = (size_t)s->ne * s->np * 4;
    /* forward: keep every stage state */
    double *st = malloc(sizeof(double) * n * 4 * (size_t)(nsteps > 0 ? nsteps : 1));
    double *U = malloc(sizeof(double) * n), *f = malloc(sizeof(double) * n), *k = malloc(sizeof(double) * n * 3);
    if (!st || !U || !f || !k) { free(st); free(U); free(f); free(k); return ORO_ERR_ARG; }
    memcpy(U, s->U, sizeof(double) * n);
    for (int a = 0; a < nsteps; ++a) {
        double *S = st + (size_t)a * 4 * n;
        memcpy(S, U, sizeof(double) * n);
        rhs_at(s, S, f);
        for (size_t m = 0; m < n; ++m) { k[m] = dt * f[m]; S[n + m] = U[m] + k[m] / 2.0; }
        rhs_at(s, S + n, f);
        for (size_t m = 0; m < n; ++m) { k[n + m] = dt * f[m]; S[2 * n + m] = U[m] + k[n + m] / 2.0; }
        rhs_at(s, S + 2 * n, f);
        for (size_t m = 0; m < n; ++m) { k[2 * n + m] = dt * f[m]; S[3 * n + m] = U[m] + k[2 * n + m]; }
        rhs_at(s, S + 3 * n, f);
        for (size_t m = 0; m < n; ++m)
            U[m] = U[m] + (1.0 / 6.0) * (k[m] + 2 * k[n + m] + 2 * k[2 * n + m] + dt * f[m]);
    }
    /* reverse */
    double *lam = malloc(sizeof(double) * n), *w = malloc(sizeof(double) * n), *nu = malloc(sizeof(double) * n);
    double *out = malloc(sizeof(double) * n);
    memcpy(lam, lamT, sizeof(double) * n);
    for (int a = nsteps - 1; a >= 0; --a) {
        double *S = st + (size_t)a * 4 * n;
        memcpy(out, lam, sizeof(double) * n);
        /* stage 4: w = lam/6 */
        for (size_t m = 0; m < n; ++m) w[m] = lam[m] / 6.0;
        oro_set_state(s, S + 3 * n);
        oro_rhs_vjp(s, w, nu);
        for (size_t m = 0; m < n; ++m) { nu[m] *= dt; out[m] += nu[m]; w[m] = lam[m] / 3.0 + nu[m]; }
        /* stage 3 (U4 = U + k3) */
        oro_set_state(s, S + 2 * n);
        oro_rhs_vjp(s, w, nu);
        for (size_t m = 0; m < n; ++m) { nu[m] *= dt; out[m] += nu[m]; w[m] = lam[m] / 3.0 + 0.5 * nu[m]; }
        /* stage 2 (U3 = U + k2/2) */
        oro_set_state(s, S + n);
        oro_rhs_vjp(s, w, nu);
        for (size_t m = 0; m < n; ++m) { nu[m] *= dt; out[m] += nu[m]; w[m] = lam[m] / 6.0 + 0.5 * nu[m]; }
        /* stage 1 (U2 = U + k1/2) */
        oro_set_state(s, S);
        oro_rhs_vjp(s, w, nu);
        for (size_t m = 0; m < n; ++m) out[m] += dt * nu[m];
        memcpy(lam, out, sizeof(double) * n);
    }
    memcpy(lam0, lam, sizeof(double) * n);
    oro_set_state(s, U);
    free(st); free(U); free(f); free(k); free(lam); free(w); free(nu); free(out);
    return 0;
}

/* ---- vortex initial state tangent (Element.cpp:134-166) ------------------ */

void oro_vortex_tangent(const oro_solver *s, double t, double epsilon, double x0, double y0, double *dU) {
    const int np = s->np, ne = s->ne;
    const double gamma = 1.4, M = 0.5, p_inf = 20.0 / 7.0;
    const double Ui = 1.0 / sqrt(2.0), Vi = 1.0 / sqrt(2.0);
    for (int e = 0; e < ne; ++e)
        for (int i = 0; i < np; ++i) {
            const double *xy = s->nodes_phys + ((size_t)e * np + i) * 2;
            for (int par = 0; par < 3; ++par) {
                dual ep = D(epsilon, par == 0), cx = D(x0, par == 1), cy = D(y0, par == 2);
                dual dx = dsub(dc(xy[0] - Ui * t), cx), dy = dsub(dc(xy[1] - Vi * t), cy);
                dual f0 = dsub(dc(1.0), dadd(dmul(dx, dx), dmul(dy, dy)));
                dual ef = D(exp(f0.v), exp(f0.v) * f0.d), eh = D(exp(0.5 * f0.v), 0.5 * exp(0.5 * f0.v) * f0.d);
                dual f1 = dsub(dc(1.0), dscale((gamma - 1) * M * M / (8.0 * M_PI * M_PI), dmul(dmul(ep, ep), ef)));
                dual f2 = dscale((Ui * Ui + Vi * Vi) / (2.0 * M_PI), dmul(ep, eh));
                double rv = pow(f1.v, 1.0 / (gamma - 1.0)), pv = p_inf * pow(f1.v, gamma / (gamma - 1.0));
                dual rho = D(rv, rv / ((gamma - 1.0) * f1.v) * f1.d);
                dual p = D(pv, pv * gamma / ((gamma - 1.0) * f1.v) * f1.d);
                dual u = dsub(dc(Ui), dmul(f2, dy)), v = dadd(dc(Vi), dmul(f2, dx));
                dual rE = dadd(dscale(1.0 / (gamma - 1.0), p), dscale(0.5, dmul(rho, dadd(dmul(u, u), dmul(v, v)))));
                double *o = dU + (((size_t)par * ne + e) * np + i) * 4;
                o[0] = rho.d;
                o[1] = dmul(rho, u).d;
                o[2] = dmul(rho, v).d;
                o[3] = rE.d;
            }
        }
}

/* ---- objective: J = sum_e detJ_e drho^T M_ref drho ------------------------ */

double oro_mass_l2_objective(const oro_solver *s, const double *rho_target, double *dJdU) {
    const int np = s->np, ne = s->ne;
    double M[MAXNP * MAXNP];
    oro_mass_matrix_ref(s->p, s->order, M);
    double J = 0.0;
    if (dJdU) memset(dJdU, 0, sizeof(double) * (size_t)ne * np * 4);
    for (int e = 0; e < ne; ++e) {
        const double det = s->geo[e * GEO_STRIDE + GEO_DET];
        double d[MAXNP];
        for (int i = 0; i < np; ++i) d[i] = s->U[((size_t)e * np + i) * 4] - rho_target[(size_t)e * np + i];
        for (int i = 0; i < np; ++i) {
            double a = 0.0;
            for (int j = 0; j < np; ++j) a += M[i * np + j] * d[j];
            J += det * d[i] * a;
            if (dJdU) dJdU[((size_t)e * np + i) * 4] = 2.0 * det * a;
        }
    }
    return J;
}
