// vt_stage.hpp -- gfx950 device code for the DG right-hand side and the
// fused RK4 stages (FP64); instantiated per quadrature size by vt_kernels_nq.hip.
//
// One thread per element.  State lives in structure-of-arrays planes
// U[(k*NP + j)*ld + e] (k = conserved variable, j = node, e = element), so
// each plane load/store of a wavefront is one contiguous 512-byte segment.
// Per element and stage the kernel
//   * loads its NP*4 state values and the P+1 edge-node values of each of its
//     three neighbours (the only nodes whose basis functions are non-zero on
//     the shared side),
//   * forms the volume term  V = Sxi G^xi + Seta G^eta  from the nodal flux
//     in contravariant form (G^xi = (y2-y0) Fx + (x0-x2) Fy, ...), which is
//     S^x Fx + S^y Fy of Element.cpp:112-113 / Evolve.cpp:254-272 regrouped,
//   * evaluates the Roe flux at the NQ Gauss points of each side
//     (Numericalflux.cpp:22-166, including its ci = c2^2 and entropy fix) and
//     subtracts the lifted flux (Evolve.cpp:225-251),
//   * applies Minv_ref / detJ (Element.cpp:101, Evolve.cpp:290-307) and
//   * performs the RK4 stage update (Timestepping.cpp:85-305) in registers.
// Reference operators and edge tables are uniform across the grid and read
// with scalar loads from one packed table.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "vt_device.hpp"
#include "vt_kernels.hpp"

namespace vt {

// Waves per SIMD the stage kernels are compiled for (p = 2 asks for 3).
constexpr int kStageMinWaves = 2;

template <int P>
__device__ __forceinline__ constexpr int edge_node(int s, int m) {
    if constexpr (P == 0) return 0;
    else if constexpr (P == 1) return s == 0 ? (m == 0 ? 0 : 1) : s == 1 ? (m == 0 ? 1 : 2) : (m == 0 ? 2 : 0);
    else if constexpr (P == 2) {
        constexpr int t[3][3] = {{0, 1, 2}, {2, 4, 5}, {5, 3, 0}};
        return t[s][m];
    } else {
        constexpr int t[3][4] = {{0, 1, 2, 3}, {3, 6, 8, 9}, {9, 7, 4, 0}};
        return t[s][m];
    }
}

// Reciprocal from the hardware estimate (v_rcp_f64) refined by two Newton
// steps (~0.5 ulp, no special-value handling: the callers' operands are
// positive normal numbers or the state is already reported non-physical).
__device__ __forceinline__ double fast_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Reciprocal square root from the hardware estimate (v_rsq_f64) refined by
// two Newton steps (positive normal operands, as fast_rcp).
__device__ __forceinline__ double fast_rsq(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// Roe flux of Numericalflux.cpp:22-166 with the reference's eigenvalue
// "sound speed" ci = c2^2 (Numericalflux.cpp:117).  1/rho_L, 1/rho_R and the
// Roe ratio sqrt(rho_R/rho_L) come from one refined 1/sqrt(rho_L rho_R)
// (within 4 ulp of the divisions, 2 ulp of the square root), the other
// divisions are refined reciprocals (1/(1+di), 1/ci, and 1/eps = 10/ci),
// the pressure as (gamma-1) rho E - (gamma-1)/2 (m_x u + m_y v), the entropy
// fix's |l| < eps as one comparison of |l|, and the flux assembly with its
// halves and common factors folded (one multiply by 1/2 per component, C1 and
// C2 sharing G1/ci); all are reassociations of the reference arithmetic at the
// ulp level.  Returns the checks' outcome as bits: 1 left state non-physical,
// 2 right state non-physical, 4 c2 <= 0 (flux_code turns them into the
// reference's error number).  This form returns twice the flux, F2 = 2 F
// (the halving is exact, so 0.5 * F2 has the bits of F): the stage kernel
// folds the 1/2 into its lift weights instead of spending four multiplies.
__device__ __forceinline__ int roe_flux2_bits(const double (&UL)[4], const double (&UR)[4], double nx, double ny,
                                              double (&F2)[4]) {
    constexpr double gmi = 0.4;
    // States: u = m/rho, p = (gamma-1)(rho E - (m_x u + m_y v)/2), H = (rho E + p)/rho.
    // One refined reciprocal square root of rho_L rho_R gives both
    // reciprocals and the Roe ratio sqrt(rho_R / rho_L).  When the densities
    // differ in sign or one is zero (a state the checks below report), each
    // side is taken on its own, so that each side's check sees its own values.
    const double rL = UL[0], rR = UR[0], prd = rL * rR;
    double irL, irR, di;
    if (prd > 0) {
        const double t = fast_rsq(prd), tt = t * t;
        irL = rR * tt;
        irR = rL * tt;
        di = rR * t;
    } else {
        irL = fast_rcp(rL);
        irR = fast_rcp(rR);
        di = sqrt(rR * irL);
    }
    const double uL = UL[1] * irL, vL = UL[2] * irL;
    const double unL = uL * nx + vL * ny;
    const double pL = fma(-0.5 * gmi, fma(UL[2], vL, UL[1] * uL), gmi * UL[3]);
    const double uR = UR[1] * irR, vR = UR[2] * irR;
    const double unR = uR * nx + vR * ny;
    const double pR = fma(-0.5 * gmi, fma(UR[2], vR, UR[1] * uR), gmi * UR[3]);
    int bits = ((pL <= 0) || (rL <= 0)) ? 1 : 0;
    if ((pR <= 0) || (rR <= 0)) bits |= 2;
    const double rHL = UL[3] + pL, HL = rHL * irL;
    const double rHR = UR[3] + pR, HR = rHR * irR;
    const double du0 = UR[0] - UL[0], du1 = UR[1] - UL[1], du2 = UR[2] - UL[2], du3 = UR[3] - UL[3];
    // Roe averages.
    const double d1 = fast_rcp(1.0 + di);
    const double ui = (di * uR + uL) * d1;
    const double vi = (di * vR + vL) * d1;
    const double Hi = (di * HR + HL) * d1;
    const double af = 0.5 * (ui * ui + vi * vi);
    const double ucp = ui * nx + vi * ny;
    const double c2 = gmi * (Hi - af);
    if (c2 <= 0) bits |= 4;
    // Eigenvalues with the reference's "sound speed" ci = c2^2 and its entropy
    // fix: |l| < eps (eps = ci/10) is replaced by (eps + l^2/eps)/2.
    const double ci = c2 * c2;
    const double ci1 = fast_rcp(ci);
    const double eps = ci * 0.1, heps = ci * 0.05, hieps = 5.0 * ci1;
    double l0 = ucp + ci, l1 = ucp - ci, l2 = ucp;
    if (fabs(l0) < eps) l0 = fma(l0 * l0, hieps, heps);
    if (fabs(l1) < eps) l1 = fma(l1 * l1, hieps, heps);
    if (fabs(l2) < eps) l2 = fma(l2 * l2, hieps, heps);
    l0 = fabs(l0); l1 = fabs(l1); l2 = fabs(l2);
    const double s1 = 0.5 * (l0 + l1);
    const double s2 = 0.5 * (l0 - l1);
    const double sl = s1 - l2;
    const double G1 = gmi * (af * du0 - ui * du1 - vi * du2 + du3);
    const double G2 = -ucp * du0 + du1 * nx + du2 * ny;
    // C1 = G1 (s1 - l2) ci1^2 + G2 s2 ci1,  C2 = G1 s2 ci1 + G2 (s1 - l2)
    const double a1 = G1 * ci1;
    const double C1 = ci1 * fma(a1, sl, G2 * s2);
    const double C2 = fma(a1, s2, G2 * sl);
    // 2 F = (F_L.n + F_R.n) - (l2 du + C1 (1, ui, vi, Hi) + C2 (0, nx, ny, ucp))
    const double pp = pL + pR;
    F2[0] = fma(rL, unL, rR * unR) - fma(l2, du0, C1);
    F2[1] = fma(UL[1], unL, fma(UR[1], unR, pp * nx)) - fma(l2, du1, fma(C1, ui, C2 * nx));
    F2[2] = fma(UL[2], unL, fma(UR[2], unR, pp * ny)) - fma(l2, du2, fma(C1, vi, C2 * ny));
    F2[3] = fma(rHL, unL, rHR * unR) - fma(l2, du3, fma(C1, Hi, C2 * ucp));
    return bits;
}

// The Roe flux F itself (roe_flux2_bits halved).
__device__ __forceinline__ int roe_flux_bits(const double (&UL)[4], const double (&UR)[4], double nx, double ny,
                                             double (&F)[4]) {
    const int bits = roe_flux2_bits(UL, UR, nx, ny, F);
#pragma unroll
    for (int k = 0; k < 4; ++k) F[k] *= 0.5;
    return bits;
}

// The reference's error number of Numericalflux.cpp:42-53, 75-78, 111-114:
// the left state is checked first, then the right one, then c2.
__device__ __forceinline__ int flux_code(int bits) {
    return (bits & 1) ? 1 : (bits & 2) ? 2 : (bits & 4) ? 3 : 0;
}

// The same outcome seen from the other side of the edge (left and right swapped).
__device__ __forceinline__ int swap_sides(int bits) {
    return ((bits & 1) << 1) | ((bits >> 1) & 1) | (bits & 4);
}

// Returns 0 or the reference's error number.
__device__ __forceinline__ int roe_flux(const double (&UL)[4], const double (&UR)[4], double nx, double ny,
                                        double (&F)[4]) {
    return flux_code(roe_flux_bits(UL, UR, nx, ny, F));
}

// Value of the neighbouring lane (lane ^ 1) of the wavefront, by a DPP
// quad permutation [1, 0, 3, 2]; both lanes of each pair must be active.
__device__ __forceinline__ int lane_pair_swap(int x) {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ double lane_pair_swap(double x) {
    const int lo = lane_pair_swap(__double2loint(x));
    const int hi = lane_pair_swap(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

// Operator tables are read through the constant address space: no kernel
// writes them, and reads at wavefront-uniform offsets then compile to scalar
// (SMEM) loads instead of per-lane vector loads with 64-bit address math.
typedef const __attribute__((address_space(4))) double* ctab_t;
__device__ __forceinline__ ctab_t const_table(const double* p) { return (ctab_t)p; }

// A zero the compiler cannot see through (a scalar register move), used to
// give table reads distinct addresses so that they are reloaded where used
// instead of being kept live in scalar registers across a long region.
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("s_mov_b32 %0, 0" : "=s"(z));
    return z;
}

__device__ __forceinline__ void report(unsigned long long* err, int stage, unsigned e, int q, int s, int code) {
    const unsigned long long key = (static_cast<unsigned long long>(stage) << 40) |
                                   ((((static_cast<unsigned long long>(e) * 16 + q) * 4 + s) * 4) + code);
    atomicMin(err, key);
}

// Plane access: element e's value of plane `plane` at byte offset
// plane * ld8 + e * 8 from base, with the element offset a 32-bit byte offset.
__device__ __forceinline__ const double& plane_at(const double* base, int plane, size_t ld8, unsigned off8) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + static_cast<size_t>(plane) * ld8 + off8);
}
__device__ __forceinline__ double& plane_at(double* base, int plane, size_t ld8, unsigned off8) {
    return *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + static_cast<size_t>(plane) * ld8 + off8);
}

// plane_at with the plane base kept in scalar registers: an empty asm hides
// base + plane * ld8 from the compiler, which otherwise folds the uniform plane
// offset into a per-lane 64-bit address (one 64-bit multiply-add per access),
// so loads take the global saddr + 32-bit voffset form (the nontemporal
// builtins still form a 64-bit address, with one 64-bit add).  base and plane
// must be wavefront-uniform; a divergent one fails to compile (the asm operand
// is a scalar register).
typedef __attribute__((address_space(1))) char* gbyte_t;
__device__ __forceinline__ const double& uplane_at(const double* base, int plane, size_t ld8, unsigned off8) {
    gbyte_t pb = (gbyte_t)(const_cast<double*>(base)) + static_cast<size_t>(plane) * ld8;
    asm("" : "+s"(pb));
    return *(const double*)(pb + off8);
}
__device__ __forceinline__ double& uplane_at(double* base, int plane, size_t ld8, unsigned off8) {
    gbyte_t pb = (gbyte_t)base + static_cast<size_t>(plane) * ld8;
    asm("" : "+s"(pb));
    return *(double*)(pb + off8);
}

// One tile of BS consecutive elements: BS threads, one element each.  `su`
// holds at least 4 * NP * BS doubles of LDS.  Every thread of the block
// calls this (it contains one barrier); threads past the mesh end only take
// part in the barrier.
template <int P, int NQ, int BS>
__device__ __forceinline__ void stage_tile(const StageParams& prm, int tile, double* su) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    constexpr int NE = P + 1;
    constexpr int NV = 4 * NP;
    constexpr double gmi = 1.4 - 1.0;
    // The block's stage state, one LDS column per thread (value v of thread t
    // at su[v * BS + t]): read once from HBM with all loads in flight
    // together, then re-read by the volume term, the traces of the element and
    // of its in-block neighbours, and the stage update.
    const int base = tile * BS;
    const int e = base + static_cast<int>(threadIdx.x);
    const bool valid = e < prm.ne;
    const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
    const unsigned oe = static_cast<unsigned>(e) * 8u;
    const ctab_t T = const_table(prm.tab);
    const double* __restrict__ Uin = prm.Uin;
    double* my = su + threadIdx.x;

    // Neighbours inside this block are read from the block's LDS copy after
    // the barrier; the others (vertical neighbours, row wrap-arounds) from
    // global memory.
    int loc[3];
    unsigned onb[3];
    double x0 = 0, y0 = 0, x1 = 0, y1 = 0, x2 = 0, y2 = 0;
    double un[NE][4];
    if (valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) my[v * BS] = uplane_at(Uin, v, ld8, oe);
        x0 = uplane_at(prm.vtx, 0, ld8, oe); y0 = uplane_at(prm.vtx, 1, ld8, oe);
        x1 = uplane_at(prm.vtx, 2, ld8, oe); y1 = uplane_at(prm.vtx, 3, ld8, oe);
        x2 = uplane_at(prm.vtx, 4, ld8, oe); y2 = uplane_at(prm.vtx, 5, ld8, oe);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const int nb = prm.nbr[s * prm.ld + e];
            loc[s] = nb - base;
            onb[s] = static_cast<unsigned>(nb) * 8u;
        }
        // Edge-node values of the neighbour across side 1, in flight during the volume term.
        if (static_cast<unsigned>(loc[0]) >= static_cast<unsigned>(BS)) {
#pragma unroll
            for (int m = 0; m < NE; ++m)
#pragma unroll
                for (int k = 0; k < 4; ++k) un[m][k] = uplane_at(Uin, k * NP + edge_node<P>(0, m), ld8, onb[0]);
        }
    }
    __syncthreads();
    if (!valid) return;
    if (static_cast<unsigned>(loc[0]) < static_cast<unsigned>(BS)) {
#pragma unroll
        for (int m = 0; m < NE; ++m)
#pragma unroll
            for (int k = 0; k < 4; ++k) un[m][k] = su[(k * NP + edge_node<P>(0, m)) * BS + loc[0]];
    }

    // Volume term: r = Sxi G^xi + Seta G^eta with G = adj(J) F (detJ * invJ = adj(J)).
    const double a00 = y2 - y0, a01 = x0 - x2, a10 = y0 - y1, a11 = x1 - x0;
    double r[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int k = 0; k < 4; ++k) r[i][k] = 0.0;
#pragma unroll 1
    for (int j = 0; j < NP; ++j) {
        const double rho = my[(0 * NP + j) * BS];
        const double m1 = my[(1 * NP + j) * BS];
        const double m2 = my[(2 * NP + j) * BS];
        const double m3 = my[(3 * NP + j) * BS];
        const double ir = fast_rcp(rho);
        const double uu = m1 * ir, vv = m2 * ir, E = m3 * ir;
        const double p = rho * gmi * (E - 0.5 * (uu * uu + vv * vv));
        const double H = E + p * ir;
        const double fx[4] = {rho * uu, rho * uu * uu + p, rho * uu * vv, rho * uu * H};
        const double fy[4] = {rho * vv, rho * uu * vv, rho * vv * vv + p, rho * vv * H};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double gx = a00 * fx[k] + a01 * fy[k];
            const double gy = a10 * fx[k] + a11 * fy[k];
            // Two fused multiply-adds into the accumulator per (i, k), instead
            // of forming the two products' sum first and adding it (a multiply,
            // a fused multiply-add and an add).
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                r[i][k] = fma(T[tab_s(P) + (j * NP + i) * 2], gx, r[i][k]);
                r[i][k] = fma(T[tab_s(P) + (j * NP + i) * 2 + 1], gy, r[i][k]);
            }
        }
    }

    // Surface term: Roe flux at the Gauss points of each side, lifted with
    // len_s * w_q * phi_i (Evolve.cpp:242-244) and subtracted (R = V - L).
    // Only the P+1 edge nodes of the element and of its neighbour enter; the
    // next side's neighbour values are requested before this side is computed.
    // Only detJ survives the volume term; each side reloads its two vertices
    // (cache hits, through an opaque offset so they are not kept live from the
    // prologue), which keeps the surface loop free of register spills.
    const double detJ = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int zv = opaque_zero();
        const double dx = uplane_at(prm.vtx, 2 * ((s + 1) % 3) + zv, ld8, oe) - uplane_at(prm.vtx, 2 * s + zv, ld8, oe);
        const double dy = uplane_at(prm.vtx, 2 * ((s + 1) % 3) + 1 + zv, ld8, oe) - uplane_at(prm.vtx, 2 * s + 1 + zv, ld8, oe);
        // |side| and the unit normal from one refined 1/|side|.
        const double l2s = dx * dx + dy * dy, il = fast_rsq(l2s);
        const double len = l2s * il, nxv = dy * il, nyv = -dx * il;
        const ctab_t tp = T + tab_tp(P) + s * NQ * NE;
        const ctab_t tm = T + tab_tm(P, NQ) + s * NQ * NE;
        const ctab_t tl = T + tab_lift(P, NQ) + s * NQ * NE;
        // The lift of 2 F with weights len/2 * w: the same products as
        // len * w * F bit for bit (halving and doubling are exact).
        const double hlen = 0.5 * len;
        auto lift = [&](const double (&tw)[NE], const double (&F2)[4]) {
#pragma unroll
            for (int m = 0; m < NE; ++m) {
                const double w = hlen * tw[m];
#pragma unroll
                for (int k = 0; k < 4; ++k) r[edge_node<P>(s, m)][k] -= w * F2[k];
            }
        };
        // On a paired mesh side s = 1 is the diagonal of the element's cell,
        // shared with lane ^ 1: each of the two evaluates the points q < ceil(NQ/2) of
        // its own orientation and takes the rest from the other, whose point
        // NQ-1-q is the same physical point with the normal reversed, by the
        // flux's antisymmetry F(a, b, n) = -F(b, a, -n).  The other lane's
        // check outcome comes along, so every element reports its own points.
        const bool half = (s == 1) && prm.pairs;
        const int nqe = half ? (NQ + 1) / 2 : NQ;
        double ctp[NE], ctm[NE];
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            ctp[m] = tp[m];
            ctm[m] = tm[m];
        }
#pragma unroll 1
        for (int q = 0; q < nqe; ++q) {
            double UL[4], UR[4], F[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                double a = 0.0, b = 0.0;
#pragma unroll
                for (int m = 0; m < NE; ++m) {
                    a += ctp[m] * my[(k * NP + edge_node<P>(s, m)) * BS];
                    b += ctm[m] * un[m][k];
                }
                UL[k] = a;
                UR[k] = b;
            }
            // Table values of the lifts of this point and of the next point's
            // traces, requested before the flux so that their scalar loads are
            // in flight during it.
            const int qn = q + 1 < nqe ? q + 1 : q;
            double ctl[NE], crl[NE];
#pragma unroll
            for (int m = 0; m < NE; ++m) {
                ctp[m] = tp[qn * NE + m];
                ctm[m] = tm[qn * NE + m];
                ctl[m] = tl[q * NE + m];
                crl[m] = tl[(NQ - 1 - q) * NE + m];
            }
            // The error number is formed only on the (rare) failing lanes:
            // flux_code is non-zero exactly when its bits are.
            const int bits = roe_flux2_bits(UL, UR, nxv, nyv, F);
            if (half && q != NQ - 1 - q) {
                double G[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) G[k] = -lane_pair_swap(F[k]);
                const int pbits = lane_pair_swap(bits);
                if (pbits) report(prm.err, prm.stage_id, prm.err_base + e, NQ - 1 - q, s, flux_code(swap_sides(pbits)));
                lift(crl, G);
            }
            if (bits) report(prm.err, prm.stage_id, prm.err_base + e, q, s, flux_code(bits));
            lift(ctl, F);
        }
        if (s < 2) {
            if (static_cast<unsigned>(loc[s + 1]) < static_cast<unsigned>(BS)) {
#pragma unroll
                for (int m = 0; m < NE; ++m)
#pragma unroll
                    for (int k = 0; k < 4; ++k) un[m][k] = su[(k * NP + edge_node<P>(s + 1, m)) * BS + loc[s + 1]];
            } else {
#pragma unroll
                for (int m = 0; m < NE; ++m)
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        un[m][k] = uplane_at(Uin, k * NP + edge_node<P>(s + 1, m), ld8, onb[s + 1]);
            }
        }
    }

    // dU/dt = Minv_ref R / detJ, then the stage update.  The update's loads and
    // stores (acc, Un, the stage output) are nontemporal: each value is touched
    // once by this kernel, and streaming them past the caches keeps the cached
    // neighbour traces of the other tiles resident.  One uniform branch
    // per mode, so that the mode's loads of acc / Un are all issued before the
    // Minv products that hide their latency.
    const double idet = 1.0 / detJ;
    const double dt = prm.dt;
    constexpr double c6 = 1.0 / 6.0;
    auto dudt_m = [&](ctab_t M, int i, int k) {
        double d = 0.0;
#pragma unroll
        for (int j = 0; j < NP; ++j) d += M[tab_minv(P) + i * NP + j] * r[j][k];
        return d * idet;
    };
    auto dudt = [&](int i, int k) { return dudt_m(T, i, k); };
    const int mode = prm.mode;
    if (mode == kModeRk2 || mode == kModeRk3) {
        // Two halves of the variables: the second half's loads are issued once
        // the first half's registers are free.
        //   stage 2: U3 = Un + k2/2;
        //   stage 3: acc = Un + (k1 + 2 k2 + 2 k3)/6 and U4 = Un + k3, with
        //            k1 = 2 (U2 - Un) and k2 = 2 (U3 - Un) recovered from the
        //            stored U2 and the staged U3, so that no stage stores k1
        //            (the differences are exact when the two states are within
        //            a factor of two, Sterbenz; k1 and k2 then carry the
        //            rounding of U2 and U3).  U2 may be read from Uout: each
        //            thread reads its own element's values before writing them.
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double ua[2 * NP], u0[2 * NP];
#pragma unroll
            for (int t = 0; t < 2 * NP; ++t) u0[t] = __builtin_nontemporal_load(&uplane_at(prm.Un, 2 * h * NP + t, ld8, oe));
            if (mode == kModeRk3) {
#pragma unroll
                for (int t = 0; t < 2 * NP; ++t) ua[t] = __builtin_nontemporal_load(&uplane_at(prm.U2, 2 * h * NP + t, ld8, oe));
            }
            // Node-major, so each row of Minv is read once per half and dies; the
            // table base is made opaque per half so that its loads are not
            // merged across the halves (that would keep all 100 entries live in
            // scalar registers and spill them).
            const ctab_t M = T + opaque_zero();
#pragma unroll
            for (int i = 0; i < NP; ++i)
#pragma unroll
                for (int kk2 = 0; kk2 < 2; ++kk2) {
                    const int k = 2 * h + kk2, t = kk2 * NP + i, v = k * NP + i;
                    const double kk = dt * dudt_m(M, i, k);
                    if (mode == kModeRk2) {
                        __builtin_nontemporal_store(u0[t] + kk * 0.5, &uplane_at(prm.Uout, v, ld8, oe));
                    } else {
                        const double k1 = 2.0 * (ua[t] - u0[t]);
                        const double k2 = 2.0 * (my[v * BS] - u0[t]);
                        __builtin_nontemporal_store(u0[t] + c6 * ((k1 + 2.0 * k2) + 2.0 * kk), &uplane_at(prm.acc, v, ld8, oe));
                        __builtin_nontemporal_store(u0[t] + kk, &uplane_at(prm.Uout, v, ld8, oe));
                    }
                }
            __builtin_amdgcn_sched_barrier(0);
        }
    } else if (mode == kModeRk4) {  // U_{n+1} = acc + k4/6
        double ac[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) ac[v] = __builtin_nontemporal_load(&uplane_at(prm.acc, v, ld8, oe));
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int v = k * NP + i;
                __builtin_nontemporal_store(ac[v] + c6 * (dt * dudt(i, k)), &uplane_at(prm.Uout, v, ld8, oe));
            }
    } else if (mode == kModeRk1) {  // k1 = dt f(Un); U2 = Un + k1/2
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int v = k * NP + i;
                const double kk = dt * dudt(i, k);
                __builtin_nontemporal_store(my[v * BS] + kk * 0.5, &uplane_at(prm.Uout, v, ld8, oe));
            }
    } else if (mode == kModeEuler) {  // U + dt dU/dt
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int v = k * NP + i;
                __builtin_nontemporal_store(my[v * BS] + dt * dudt(i, k), &uplane_at(prm.Uout, v, ld8, oe));
            }
    } else {  // kModeRhs
#pragma unroll
        for (int i = 0; i < NP; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(dudt(i, k), &uplane_at(prm.Uout, k * NP + i, ld8, oe));
    }
}


// XCD-aware tile order: consecutive workgroups are dispatched round-robin
// over the 8 XCDs, so workgroup b works on the band of tiles
// [(b % 8) * per_xcd, (b % 8 + 1) * per_xcd) of its XCD.  Each XCD then sweeps
// one contiguous band of elements, and the vertical neighbours (one element
// row away) of the band's rows are read through that XCD's own L2.  Where a
// mesh row holds a multiple of 8 tiles, plain order (workgroup b on tile b,
// so tile t on XCD t % 8) keeps the neighbours one row away on the same XCD
// as well and is about 1% faster; the context chooses it then (plain_order).
// Occupancy: p = 2 is held to 168 VGPRs for 3 waves per SIMD (a few spills,
// 2.5% faster than 2 waves); p = 0, 1 reach 3 or more waves as compiled, p = 3
// runs 2 waves (232 VGPRs).
template <int P, int NQ, int BS>
__global__ void __launch_bounds__(BS, (P == 2 ? 3 : kStageMinWaves)) stage_kernel(StageParams prm) {
    __shared__ double su[4 * ((P + 1) * (P + 2) / 2) * BS];
    const int b = blockIdx.x;
    stage_tile<P, NQ, BS>(prm, prm.plain_order ? b : (b & 7) * prm.blocks_per_xcd + (b >> 3), su);
}

// Persistent form: gridDim.x (a multiple of 8) resident workgroups, each
// looping over the tiles of its XCD's band.
template <int P, int NQ, int BS>
__global__ void __launch_bounds__(BS, kStageMinWaves) stage_kernel_persistent(StageParams prm) {
    __shared__ double su[4 * ((P + 1) * (P + 2) / 2) * BS];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, slots = gridDim.x >> 3;
    for (int t = slot; t < prm.blocks_per_xcd; t += slots) {
        stage_tile<P, NQ, BS>(prm, xcd * prm.blocks_per_xcd + t, su);
        __syncthreads();
    }
}

// Ensemble form: the tiles of up to kMaxGroups element groups (instances
// packed by polynomial order) in one launch of 8 * slots workgroups; each
// workgroup takes items slot, slot + slots, ... of its XCD's walk.  With
// slots = kbegin[ngroups] every workgroup does one tile.  The tile's group
// selects the P code path (uniform over the workgroup).
template <int NQ, int BS>
__global__ void __launch_bounds__(BS, kStageMinWaves) batched_stage_kernel(BatchParams bp) {
    __shared__ double su[4 * 10 * BS];
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, slots = gridDim.x >> 3;
    for (int k = slot; k < bp.kbegin[bp.ngroups]; k += slots) {
        int g = 0;
#pragma unroll
        for (int i = 1; i < kMaxGroups; ++i)
            if (i < bp.ngroups && k >= bp.kbegin[i]) g = i;
        const int lt = xcd * bp.tpx[g] + (k - bp.kbegin[g]);
        if (lt < bp.tiles[g]) {
            switch (bp.p[g]) {
            case 0: stage_tile<0, NQ, BS>(bp.g[g], lt, su); break;
            case 1: stage_tile<1, NQ, BS>(bp.g[g], lt, su); break;
            case 2: stage_tile<2, NQ, BS>(bp.g[g], lt, su); break;
            default: stage_tile<3, NQ, BS>(bp.g[g], lt, su); break;
            }
        }
        __syncthreads();
    }
}

// Host-side launchers of one quadrature size (defined in vt_kernels_nq.hip).
template <int NQ>
struct NqLaunch {
    static hipError_t stage(const StageParams& prm, int p, const LaunchConfig& cfg, hipStream_t st);
    static hipError_t batched(const BatchParams& bp, const LaunchConfig& cfg, hipStream_t st);
    static hipError_t adjoint(const AdjParams& prm, int p, hipStream_t st);
    static hipError_t small(const SmallParams& prm, int p, hipStream_t st);
};

template <int NQ, int BS, int P>
hipError_t launch_one(const StageParams& prm, const LaunchConfig& cfg, hipStream_t st) {
    if (cfg.persistent && cfg.slots_per_xcd > 0 && cfg.slots_per_xcd < prm.blocks_per_xcd) {
        hipLaunchKernelGGL((stage_kernel_persistent<P, NQ, BS>), dim3(8 * cfg.slots_per_xcd), dim3(BS), 0, st, prm);
    } else {
        hipLaunchKernelGGL((stage_kernel<P, NQ, BS>), dim3(8 * prm.blocks_per_xcd), dim3(BS), 0, st, prm);
    }
    return hipGetLastError();
}

template <int NQ, int BS>
hipError_t launch_bs(const StageParams& prm, int p, const LaunchConfig& cfg, hipStream_t st) {
    switch (p) {
    case 0: return launch_one<NQ, BS, 0>(prm, cfg, st);
    case 1: return launch_one<NQ, BS, 1>(prm, cfg, st);
    case 2: return launch_one<NQ, BS, 2>(prm, cfg, st);
    case 3: return launch_one<NQ, BS, 3>(prm, cfg, st);
    default: return hipErrorInvalidValue;
    }
}

template <int NQ>
hipError_t NqLaunch<NQ>::stage(const StageParams& prm, int p, const LaunchConfig& cfg, hipStream_t st) {
    if (cfg.block == 64) return launch_bs<NQ, 64>(prm, p, cfg, st);
    return cfg.block == 128 ? launch_bs<NQ, 128>(prm, p, cfg, st) : launch_bs<NQ, 256>(prm, p, cfg, st);
}

template <int NQ>
hipError_t NqLaunch<NQ>::batched(const BatchParams& bp, const LaunchConfig& cfg, hipStream_t st) {
    const int per_xcd = bp.kbegin[bp.ngroups];
    const int slots = cfg.persistent && cfg.slots_per_xcd > 0 && cfg.slots_per_xcd < per_xcd ? cfg.slots_per_xcd : per_xcd;
    if (slots < 1) return hipSuccess;
    if (cfg.block == 64)
        hipLaunchKernelGGL((batched_stage_kernel<NQ, 64>), dim3(8 * slots), dim3(64), 0, st, bp);
    else if (cfg.block == 128)
        hipLaunchKernelGGL((batched_stage_kernel<NQ, 128>), dim3(8 * slots), dim3(128), 0, st, bp);
    else
        hipLaunchKernelGGL((batched_stage_kernel<NQ, 256>), dim3(8 * slots), dim3(256), 0, st, bp);
    return hipGetLastError();
}

}  // namespace vt
