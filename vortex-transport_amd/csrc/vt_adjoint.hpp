// vt_adjoint.hpp -- gfx950 device code of the discrete adjoint of the DG
// right-hand side and of the RK4 step (FP64); instantiated per quadrature size
// by vt_kernels_nq.hip next to the forward stage kernels.
//
// The forward map per element is f_e(U) = Minv_ref (V_e - L_e) / detJ_e with
// the volume term V_e = Sxi G^xi + Seta G^eta (G = adj(J) F(U_j)) and the
// lifted Roe flux L_e.  For a cotangent w the adjoint g = f'(U)^T w is formed
// element by element in gather form, so no element writes to another:
//   y_e   = Minv_ref^T w_e / detJ_e                    (input plane set Y)
//   volume: z = Sxi^T y, Seta^T y; bFx = a00 z^xi + a10 z^eta,
//           bFy = a01 z^xi + a11 z^eta; g_j += (dF/dU_j)^T (bFx, bFy);
//   surface: at each Gauss point q of side s the element's own flux
//           F(UL, UR, n) receives the cotangent c_e(q) = len sum_m lift[s,q,m] y_e,m
//           and the neighbour's flux at the same physical point (its point
//           q~ = NQ-1-q, same side index, normal -n) receives c_nb(q~).  The
//           Roe flux is antisymmetric, F(a, b, n) = -F(b, a, -n), so the
//           neighbour's derivative with respect to this element's trace is
//           -(dF/dUL) at this element's point, and both contributions fold into
//           one reverse sweep:  g_{edge(s,m)} -= tp[s,q,m] (dF/dUL)^T (c_e(q) - c_nb(q~)).
// The reverse RK4 step (stage states S0..S3 of the step, lam = dJ/dU_{n+1}):
//   nu4 = dt f'(S3)^T (lam/6);          acc = lam + nu4
//   nu3 = dt f'(S2)^T (lam/3 + nu4);    acc += nu3
//   nu2 = dt f'(S1)^T (lam/3 + nu3/2);  acc += nu2
//   nu1 = dt f'(S0)^T (lam/6 + nu2/2);  lam_n = acc + nu1
// one kernel per stage; each also writes y of the next stage's cotangent.
#pragma once
#include <hip/hip_runtime.h>

#include "vt_stage.hpp"

namespace vt {

// (dF/dUL)^T bF, and with RIGHT also (dF/dUR)^T bF, for the Roe flux of
// Numericalflux.cpp:31-156, differentiated on the branches the forward takes
// (entropy fix, |.|); reciprocals as in roe_flux.  The cotangent comes halved,
// hF = bF / 2: every path from F to the state carries the flux's factor 1/2,
// which the caller folds into its weights.
template <bool RIGHT>
__device__ __forceinline__ void roe_flux_vjp(const double (&UL)[4], const double (&UR)[4], double nx, double ny,
                                             const double (&hF)[4], double (&bUL)[4], double (&bUR)[4]) {
    constexpr double gmi = 0.4;
    // forward sweep
    const double rL = UL[0], irL = fast_rcp(rL);
    const double uL = UL[1] * irL, vL = UL[2] * irL, unL = uL * nx + vL * ny;
    const double pL = gmi * (UL[3] - 0.5 * (UL[1] * UL[1] + UL[2] * UL[2]) * irL);
    const double rHL = UL[3] + pL, HL = rHL * irL;
    const double rR = UR[0], irR = fast_rcp(rR);
    const double uR = UR[1] * irR, vR = UR[2] * irR;
    const double pR = gmi * (UR[3] - 0.5 * (UR[1] * UR[1] + UR[2] * UR[2]) * irR);
    const double HR = (UR[3] + pR) * irR;
    const double du0 = UR[0] - UL[0], du1 = UR[1] - UL[1], du2 = UR[2] - UL[2], du3 = UR[3] - UL[3];
    const double di = sqrt(rR * irL);
    const double d1 = fast_rcp(1.0 + di);
    const double ui = (di * uR + uL) * d1, vi = (di * vR + vL) * d1, Hi = (di * HR + HL) * d1;
    const double af = 0.5 * (ui * ui + vi * vi);
    const double ucp = ui * nx + vi * ny;
    const double c2 = gmi * (Hi - af);
    const double ci = c2 * c2, ci1 = fast_rcp(ci), eps = 0.1 * ci, ieps = 10.0 * ci1;
    const double lr0 = ucp + ci, lr1 = ucp - ci, lr2 = ucp;
    const bool f0 = (lr0 < eps) && (lr0 > -eps), f1 = (lr1 < eps) && (lr1 > -eps), f2 = (lr2 < eps) && (lr2 > -eps);
    const double lf0 = f0 ? 0.5 * (eps + lr0 * lr0 * ieps) : lr0;
    const double lf1 = f1 ? 0.5 * (eps + lr1 * lr1 * ieps) : lr1;
    const double lf2 = f2 ? 0.5 * (eps + lr2 * lr2 * ieps) : lr2;
    const double la0 = fabs(lf0), la1 = fabs(lf1), l3 = fabs(lf2);
    const double s1 = 0.5 * (la0 + la1), s2 = 0.5 * (la0 - la1), sl = s1 - l3;
    const double G1 = gmi * (af * du0 - ui * du1 - vi * du2 + du3);
    const double G2 = -ucp * du0 + du1 * nx + du2 * ny;
    const double C1 = G1 * sl * ci1 * ci1 + G2 * s2 * ci1;
    const double C2 = G1 * s2 * ci1 + G2 * sl;
    // reverse sweep
    const double bl3a = -(hF[0] * du0 + hF[1] * du1 + hF[2] * du2 + hF[3] * du3);
    double bdu0 = -l3 * hF[0], bdu1 = -l3 * hF[1], bdu2 = -l3 * hF[2], bdu3 = -l3 * hF[3];
    const double bC1 = -(hF[0] + hF[1] * ui + hF[2] * vi + hF[3] * Hi);
    const double bC2 = -(hF[1] * nx + hF[2] * ny + hF[3] * ucp);
    double bui = -C1 * hF[1], bvi = -C1 * hF[2], bHi = -C1 * hF[3], bucp = -C2 * hF[3];
    // C2 = G1 s2 ci1 + G2 (s1 - l3);  C1 = G1 (s1 - l3) ci1^2 + G2 s2 ci1
    const double bG1 = bC2 * s2 * ci1 + bC1 * sl * ci1 * ci1;
    const double bG2 = bC2 * sl + bC1 * s2 * ci1;
    const double bs1 = bC2 * G2 + bC1 * G1 * ci1 * ci1;
    const double bs2 = bC2 * G1 * ci1 + bC1 * G2 * ci1;
    const double bl3 = bl3a - bs1;
    const double bci1 = bC2 * G1 * s2 + bC1 * (2.0 * G1 * sl * ci1 + G2 * s2);
    // G2 = -ucp du0 + du1 nx + du2 ny
    bucp -= bG2 * du0;
    bdu0 -= bG2 * ucp;
    bdu1 += bG2 * nx;
    bdu2 += bG2 * ny;
    // G1 = gmi (af du0 - ui du1 - vi du2 + du3)
    const double gG1 = gmi * bG1;
    double baf = gG1 * du0;
    bdu0 += gG1 * af;
    bui -= gG1 * du1;
    bdu1 -= gG1 * ui;
    bvi -= gG1 * du2;
    bdu2 -= gG1 * vi;
    bdu3 += gG1;
    // |l| after the entropy fix; s1 = (l0 + l1)/2, s2 = (l0 - l1)/2
    const double ba0 = 0.5 * (bs1 + bs2), ba1 = 0.5 * (bs1 - bs2);
    const double bf0 = lf0 < 0 ? -ba0 : ba0, bf1 = lf1 < 0 ? -ba1 : ba1, bf2 = lf2 < 0 ? -bl3 : bl3;
    double beps = 0.0;
    double bl0 = bf0, bl1 = bf1, bl2 = bf2;
    if (f0) { bl0 = bf0 * lr0 * ieps; beps += bf0 * (0.5 - 0.5 * lr0 * lr0 * ieps * ieps); }
    if (f1) { bl1 = bf1 * lr1 * ieps; beps += bf1 * (0.5 - 0.5 * lr1 * lr1 * ieps * ieps); }
    if (f2) { bl2 = bf2 * lr2 * ieps; beps += bf2 * (0.5 - 0.5 * lr2 * lr2 * ieps * ieps); }
    bucp += bl0 + bl1 + bl2;
    const double bci = bl0 - bl1 + 0.1 * beps - bci1 * ci1 * ci1;
    const double bc2 = 2.0 * c2 * bci;
    bHi += gmi * bc2;
    baf -= gmi * bc2;
    bui += bucp * nx + baf * ui;
    bvi += bucp * ny + baf * vi;
    // Roe averages: only the left-state dependences
    const double buL = bui * d1, bvL = bvi * d1, bHL = bHi * d1;
    const double bd1 = bui * (di * uR + uL) + bvi * (di * vR + vL) + bHi * (di * HR + HL);
    const double bdi = (bui * uR + bvi * vR + bHi * HR) * d1 - bd1 * d1 * d1;
    double brL = -0.5 * bdi * di * irL;
    // left-state flux F_L = (rL unL, UL1 unL + pL nx, UL2 unL + pL ny, rHL unL), weight 1/2
    const double h0 = hF[0], h1 = hF[1], h2 = hF[2], h3 = hF[3];
    brL += h0 * unL;
    const double bun = h0 * rL + h1 * UL[1] + h2 * UL[2] + h3 * rHL;
    double bU1 = h1 * unL, bU2 = h2 * unL;
    double bp = h1 * nx + h2 * ny;
    double brH = h3 * unL + bHL * irL;
    brL -= bHL * HL * irL;
    double bU3 = brH;
    bp += brH;
    const double bu = buL + bun * nx, bv = bvL + bun * ny;
    bU3 += gmi * bp;
    bU1 -= gmi * uL * bp;
    bU2 -= gmi * vL * bp;
    brL += gmi * 0.5 * (uL * uL + vL * vL) * bp;
    bU1 += bu * irL;
    brL -= bu * uL * irL;
    bU2 += bv * irL;
    brL -= bv * vL * irL;
    bUL[0] = brL - bdu0;
    bUL[1] = bU1 - bdu1;
    bUL[2] = bU2 - bdu2;
    bUL[3] = bU3 - bdu3;
    if (RIGHT) {
        // Roe averages, right-state dependences: d(ui)/d(uR) = di d1, and
        // di = sqrt(rR / rL) gives d(di)/d(rR) = di / (2 rR).
        const double dd = di * d1;
        const double buR = bui * dd, bvR = bvi * dd, bHR = bHi * dd;
        double brR = 0.5 * bdi * di * irR;
        // right-state flux F_R = (rR unR, UR1 unR + pR nx, UR2 unR + pR ny, rHR unR), weight 1/2
        const double unR = uR * nx + vR * ny, rHR = UR[3] + pR;
        brR += h0 * unR;
        const double bunR = h0 * rR + h1 * UR[1] + h2 * UR[2] + h3 * rHR;
        double bV1 = h1 * unR, bV2 = h2 * unR;
        double bpR = h1 * nx + h2 * ny;
        const double brHR = h3 * unR + bHR * irR;
        brR -= bHR * HR * irR;
        double bV3 = brHR;
        bpR += brHR;
        const double buRt = buR + bunR * nx, bvRt = bvR + bunR * ny;
        bV3 += gmi * bpR;
        bV1 -= gmi * uR * bpR;
        bV2 -= gmi * vR * bpR;
        brR += gmi * 0.5 * (uR * uR + vR * vR) * bpR;
        bV1 += buRt * irR;
        brR -= buRt * uR * irR;
        bV2 += bvRt * irR;
        brR -= bvRt * vR * irR;
        bUR[0] = brR + bdu0;
        bUR[1] = bV1 + bdu1;
        bUR[2] = bV2 + bdu2;
        bUR[3] = bV3 + bdu3;
    }
}

// (dFx/dU)^T bFx + (dFy/dU)^T bFy of the nodal Euler flux (Timestepping.cpp:59-76),
// accumulated into g.
__device__ __forceinline__ void euler_flux_vjp(double rho, double m1, double m2, double E, const double (&bFx)[4],
                                               const double (&bFy)[4], double& g0, double& g1, double& g2,
                                               double& g3) {
    constexpr double gmi = 0.4;
    const double ir = fast_rcp(rho);
    const double u = m1 * ir, v = m2 * ir;
    const double p = gmi * (E - 0.5 * (m1 * u + m2 * v));
    const double Ep = E + p;
    double bm1 = bFx[0] + bFx[1] * u + bFx[2] * v;
    double bm2 = bFy[0] + bFy[1] * u + bFy[2] * v;
    const double bu = bFx[1] * m1 + bFx[3] * Ep + bFy[1] * m2;
    const double bv = bFx[2] * m1 + bFy[2] * m2 + bFy[3] * Ep;
    const double bp = bFx[1] + bFx[3] * u + bFy[2] + bFy[3] * v;
    const double bE = bFx[3] * u + bFy[3] * v + gmi * bp;
    bm1 += (bu - gmi * m1 * bp) * ir;
    bm2 += (bv - gmi * m2 * bp) * ir;
    const double brho = gmi * 0.5 * (u * u + v * v) * bp - (bu * u + bv * v) * ir;
    g0 += brho;
    g1 += bm1;
    g2 += bm2;
    g3 += bE;
}

// y = scale * Minv_ref^T W / detJ per element (the first cotangent of a
// reverse sweep, or of a single right-hand-side adjoint).
template <int P>
__global__ void __launch_bounds__(256) adjoint_seed_kernel(AdjParams prm) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= prm.ne) return;
    const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
    const unsigned oe = static_cast<unsigned>(e) * 8u;
    const ctab_t T = const_table(prm.tab);
    const double x0 = plane_at(prm.vtx, 0, ld8, oe), y0 = plane_at(prm.vtx, 1, ld8, oe);
    const double x1 = plane_at(prm.vtx, 2, ld8, oe), y1 = plane_at(prm.vtx, 3, ld8, oe);
    const double x2 = plane_at(prm.vtx, 4, ld8, oe), y2 = plane_at(prm.vtx, 5, ld8, oe);
    const double s = prm.scale / ((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double w[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) w[i] = plane_at(prm.lam, k * NP + i, ld8, oe);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            double a = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i) a += T[tab_minvt(P) + j * NP + i] * w[i];
            plane_at(prm.Yout, k * NP + j, ld8, oe) = s * a;
        }
    }
}

// One adjoint stage: g = f'(U)^T w for all elements, w given through y = Y,
// then the reverse-RK4 bookkeeping of the mode (kAdjRaw: Yout = g).
template <int P, int NQ>
__global__ void __launch_bounds__(kAdjBlock, 1) adjoint_stage_kernel(AdjParams prm) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    constexpr int NE = P + 1;
    constexpr int NV = 4 * NP;
    constexpr int BS = kAdjBlock;
    // The block's y (value v of thread t at sy[v * BS + t]): read once from
    // HBM, then by the element's own volume and surface terms and by the
    // surface terms of in-block neighbours.
    __shared__ double sy[NV * BS];
    const int b = blockIdx.x;
    const int base = ((b & 7) * prm.blocks_per_xcd + (b >> 3)) * BS;
    const int e = base + static_cast<int>(threadIdx.x);
    const bool valid = e < prm.ne;
    const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
    const unsigned oe = static_cast<unsigned>(e) * 8u;
    const ctab_t T = const_table(prm.tab);
    const double* __restrict__ U = prm.U;
    const double* __restrict__ Y = prm.Y;
    double* my = sy + threadIdx.x;
    double x0 = 0, y0 = 0, x1 = 0, y1 = 0, x2 = 0, y2 = 0;
    int loc[3];
    unsigned onb[3];
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) my[v * BS] = plane_at(Y, v, ld8, oe);
        x0 = plane_at(prm.vtx, 0, ld8, oe); y0 = plane_at(prm.vtx, 1, ld8, oe);
        x1 = plane_at(prm.vtx, 2, ld8, oe); y1 = plane_at(prm.vtx, 3, ld8, oe);
        x2 = plane_at(prm.vtx, 4, ld8, oe); y2 = plane_at(prm.vtx, 5, ld8, oe);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int nb = prm.nbr[s * prm.ld + e];
            loc[s] = nb - base;
            onb[s] = static_cast<unsigned>(nb) * 8u;
        }
    }
    __syncthreads();
    if (!valid) return;

    double g[NP][4];
    // Volume term.
    const double a00 = y2 - y0, a01 = x0 - x2, a10 = y0 - y1, a11 = x1 - x0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        double bFx[4], bFy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double zx = 0.0, zy = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const double yv = my[(k * NP + i) * BS];
                zx += T[tab_s(P) + (j * NP + i) * 2] * yv;
                zy += T[tab_s(P) + (j * NP + i) * 2 + 1] * yv;
            }
            bFx[k] = a00 * zx + a10 * zy;
            bFy[k] = a01 * zx + a11 * zy;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) g[j][k] = 0.0;
        euler_flux_vjp(plane_at(U, 0 * NP + j, ld8, oe), plane_at(U, 1 * NP + j, ld8, oe),
                       plane_at(U, 2 * NP + j, ld8, oe), plane_at(U, 3 * NP + j, ld8, oe), bFx, bFy, g[j][0], g[j][1],
                       g[j][2], g[j][3]);
    }

    // Surface term.
    const double vx[3] = {x0, x1, x2}, vy[3] = {y0, y1, y2};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const double dx = vx[(s + 1) % 3] - vx[s], dy = vy[(s + 1) % 3] - vy[s];
        // |side| and the unit normal from one refined 1/|side|.
        const double l2s = dx * dx + dy * dy, il = fast_rsq(l2s);
        const double hlen = 0.5 * (l2s * il), nxv = dy * il, nyv = -dx * il;
        const ctab_t tp = T + tab_tp(P) + s * NQ * NE;
        const ctab_t tm = T + tab_tm(P, NQ) + s * NQ * NE;
        const ctab_t tl = T + tab_lift(P, NQ) + s * NQ * NE;
        double yn[NE][4], un[NE][4], uo[NE][4];
        const bool inblk = static_cast<unsigned>(loc[s]) < static_cast<unsigned>(BS);
#pragma unroll
        for (int m = 0; m < NE; ++m)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int v = k * NP + edge_node<P>(s, m);
                yn[m][k] = inblk ? sy[v * BS + loc[s]] : plane_at(Y, v, ld8, onb[s]);
                un[m][k] = plane_at(U, v, ld8, onb[s]);
                uo[m][k] = plane_at(U, v, ld8, oe);
            }
        // On a paired mesh side s = 1 is the diagonal of the element's cell,
        // shared with lane ^ 1 (as in the forward stage kernel): each of the two
        // sweeps the points q < ceil(NQ/2) of its own orientation and, except at
        // a middle point, also returns the derivative with respect to the other
        // element's trace, which is the other element's own derivative at its
        // point NQ-1-q (d and F both change sign there), passed by a lane swap.
        const bool half = (s == 1) && prm.pairs;
        const int nqe = half ? (NQ + 1) / 2 : NQ;
#pragma unroll 1
        for (int q = 0; q < nqe; ++q) {
            // d(q) = len (c_e(q) - c_nb(q~)) / 2, half the cotangent of this
            // element's flux minus that of the neighbour's flux at the same point.
            const int qn = NQ - 1 - q;
            double UL[4], UR[4], d[4], bUL[4], bUR[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double a = 0.0, c = 0.0, ce = 0.0, cn = 0.0;
#pragma unroll
                for (int m = 0; m < NE; ++m) {
                    a += tp[q * NE + m] * uo[m][k];
                    c += tm[q * NE + m] * un[m][k];
                    ce += tl[q * NE + m] * my[(k * NP + edge_node<P>(s, m)) * BS];
                    cn += tl[qn * NE + m] * yn[m][k];
                }
                UL[k] = a;
                UR[k] = c;
                d[k] = hlen * (ce - cn);
            }
            if (half && q != qn) {
                roe_flux_vjp<true>(UL, UR, nxv, nyv, d, bUL, bUR);
                double bP[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) bP[k] = lane_pair_swap(bUR[k]);
#pragma unroll
                for (int m = 0; m < NE; ++m)
#pragma unroll
                    for (int k = 0; k < 4; ++k) g[edge_node<P>(s, m)][k] -= tp[qn * NE + m] * bP[k];
            } else {
                roe_flux_vjp<false>(UL, UR, nxv, nyv, d, bUL, bUR);
            }
#pragma unroll
            for (int m = 0; m < NE; ++m)
#pragma unroll
                for (int k = 0; k < 4; ++k) g[edge_node<P>(s, m)][k] -= tp[q * NE + m] * bUL[k];
        }
    }

    const int mode = prm.mode;
    if (mode == kAdjRaw) {
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) plane_at(prm.Yout, k * NP + i, ld8, oe) = g[i][k];
        return;
    }
    // Reverse-RK4 bookkeeping: the next cotangent w for all variables first
    // (g dies as w is formed), then y = Minv_ref^T w / detJ column by column.
    // Each mode loads all of its operands before its first store, so the loads
    // are in flight together (a store may alias a later load otherwise).
    const double dt = prm.dt;
    const double idet = 1.0 / ((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0));
    double w[NP][4];
    if (mode == kAdj4) {
        double lam[NP][4];
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) lam[i][k] = plane_at(prm.lam, k * NP + i, ld8, oe);
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double nu = dt * g[i][k];
                plane_at(prm.acc, k * NP + i, ld8, oe) = lam[i][k] + nu;
                w[i][k] = lam[i][k] * (1.0 / 3.0) + nu;
            }
    } else if (mode == kAdj1) {  // lam_n = acc + nu1, then the next step's first cotangent lam_n / 6
        double a[NP][4];
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) a[i][k] = plane_at(prm.acc, k * NP + i, ld8, oe);
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double ln = a[i][k] + dt * g[i][k];
                plane_at(prm.lam, k * NP + i, ld8, oe) = ln;
                w[i][k] = ln * (1.0 / 6.0);
            }
    } else {  // kAdj3, kAdj2
        const double cl = mode == kAdj3 ? (1.0 / 3.0) : (1.0 / 6.0);
        double lam[NP][4], a[NP][4];
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                lam[i][k] = plane_at(prm.lam, k * NP + i, ld8, oe);
                a[i][k] = plane_at(prm.acc, k * NP + i, ld8, oe);
            }
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const double nu = dt * g[i][k];
                plane_at(prm.acc, k * NP + i, ld8, oe) = a[i][k] + nu;
                w[i][k] = lam[i][k] * cl + 0.5 * nu;
            }
    }
#pragma unroll
    for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double a = 0.0;
#pragma unroll
            for (int i = 0; i < NP; ++i) a += T[tab_minvt(P) + j * NP + i] * w[i][k];
            plane_at(prm.Yout, k * NP + j, ld8, oe) = idet * a;
        }
}

template <int NQ, int P>
hipError_t launch_adjoint_p(const AdjParams& prm, hipStream_t st) {
    if (prm.mode == kAdjSeed) {
        hipLaunchKernelGGL((adjoint_seed_kernel<P>), dim3((prm.ne + 255) / 256), dim3(256), 0, st, prm);
    } else {
        hipLaunchKernelGGL((adjoint_stage_kernel<P, NQ>), dim3(8 * prm.blocks_per_xcd), dim3(kAdjBlock), 0, st, prm);
    }
    return hipGetLastError();
}

template <int NQ>
hipError_t NqLaunch<NQ>::adjoint(const AdjParams& prm, int p, hipStream_t st) {
    switch (p) {
    case 0: return launch_adjoint_p<NQ, 0>(prm, st);
    case 1: return launch_adjoint_p<NQ, 1>(prm, st);
    case 2: return launch_adjoint_p<NQ, 2>(prm, st);
    case 3: return launch_adjoint_p<NQ, 3>(prm, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace vt
