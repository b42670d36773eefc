// vt_small.hpp -- gfx950 device code of the small-mesh path: a whole mesh of
// at most kSmallMaxElements elements advanced by many RK4 / Euler steps in one
// launch of one workgroup, its stage state resident in LDS.
//
// A small mesh (the reference's own 8 x 8-cell case has 128 elements) leaves
// the one-thread-per-element stage kernel with two workgroups on a 256-CU
// device, each thread running its element's whole serial chain and every
// stage paying a launch.  Here one workgroup of kSmallThreads threads (16
// wavefronts on one CU, four per SIMD) holds the stage state in LDS across all
// the steps of a launch and spreads each stage over the threads in two roles
// that meet at the stage's two barriers:
//   A  threads [kSmallItems, kSmallThreads), one (shared side, Gauss point) at
//      a time: the two traces and one Roe flux evaluation, stored for both
//      elements (F for the element whose orientation was used, -F at the
//      mirrored point for its neighbour, by the flux's antisymmetry
//      F(a, b, n) = -F(b, a, -n)), so every physical flux point is evaluated
//      once instead of once per side;
//   B  threads [0, kSmallItems), one (variable k, element e) each: the volume
//      term of variable k (Sxi G^xi + Seta G^eta from the nodal flux) beside
//      the A threads' fluxes, then after the barrier the lift of the stored
//      fluxes, Minv_ref / detJ and the RK update.  U_n and the running sum
//      k1 + 2 k2 + 2 k3 stay in this thread's registers for the whole launch.
// The roles are uniform over a wavefront and each has its own loop over the
// stages, so only one role's code runs on a wavefront and each role's
// registers are its own (at most 128 per thread for four wavefronts per SIMD).
// The stage update follows Timestepping.cpp:85-305 operation by operation:
// stage states U_n + k_s / 2, U_n + k_3, and U_n + (1/6)(k1 + 2 k2 + 2 k3 + k4).
#pragma once
#include <hip/hip_runtime.h>

#include "vt_stage.hpp"

namespace vt {

// LDS layout (offsets in doubles), nep = ne rounded up to 64, rows of ns =
// nep + 1 doubles (an odd stride spreads the rows of one element over the LDS
// banks), nfp = 3 nep / 2:
//   S   [4 * NP][ns]         stage state, value v of element e at S[v * ns + e]
//   Fs  [3 * NQ * 4][ns]     flux of side s, point q, variable k at Fs[((s * NQ + q) * 4 + k) * ns + e]
//   Nf  [2][nfp]             unit normal of each listed side (of its first element)
//   Tw  [2][3][NQ][P+1]      own-side and neighbour-side trace weights
//   Fc  [nfp] ints           the side list
//   key                      first failure of the launch (error key)
struct SmallLayout {
    int nep, ns, nfp, S, Fs, Nf, Tw, Fc, key, total;
};
__host__ __device__ constexpr int small_pad(int ne) { return (ne + 63) / 64 * 64; }
__host__ __device__ inline SmallLayout small_layout(int np, int ne1, int nq, int nep) {
    SmallLayout L;
    L.nep = nep;
    L.ns = L.nep + 1;
    L.nfp = 3 * L.nep / 2;
    L.S = 0;
    L.Fs = 4 * np * L.ns;
    L.Nf = L.Fs + 12 * nq * L.ns;
    L.Tw = L.Nf + 2 * L.nfp;
    L.Fc = L.Tw + 6 * nq * ne1;
    L.key = L.Fc + (L.nfp + 1) / 2;
    L.total = L.key + 1;
    return L;
}

// Phase A item: side SD of the listed side f (elements e < nb), Gauss point q
// in e's orientation.  Traces from the stage state, one Roe flux, stored for e
// at q and, negated, for nb at NQ - 1 - q; failures of either element go to key.
template <int P, int NQ, int SD, int ns>
__device__ __forceinline__ void small_side_flux(const double* S, double* Fs, const double* Nf, const double* Tw,
                                                const int* Fc, unsigned long long* key, int nfp, int f, int q,
                                                int st) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    constexpr int NE = P + 1;
    const int fc = Fc[f];
    const int e = fc & 1023, nb = (fc >> 10) & 1023;
    const double* tp = Tw + (SD * NQ + q) * NE;
    const double* tm = tp + 3 * NQ * NE;
    double UL[4], UR[4], F[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int m = 0; m < NE; ++m) {
            a += tp[m] * S[(k * NP + edge_node<P>(SD, m)) * ns + e];
            b += tm[m] * S[(k * NP + edge_node<P>(SD, m)) * ns + nb];
        }
        UL[k] = a;
        UR[k] = b;
    }
    const int bits = roe_flux_bits(UL, UR, Nf[f], Nf[nfp + f], F);
    const int qm = NQ - 1 - q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        Fs[((SD * NQ + q) * 4 + k) * ns + e] = F[k];
        Fs[((SD * NQ + qm) * 4 + k) * ns + nb] = -F[k];
    }
    if (bits) {
        const unsigned long long stage = static_cast<unsigned long long>(st) << 40;
        const int ce = flux_code(bits), cn = flux_code(swap_sides(bits));
        if (ce) atomicMin(key, stage | ((((static_cast<unsigned long long>(e) * 16 + q) * 4 + SD) * 4) + ce));
        if (cn) atomicMin(key, stage | ((((static_cast<unsigned long long>(nb) * 16 + qm) * 4 + SD) * 4) + cn));
    }
}

// Volume term of variable K at element e: r_i += Sxi_ij G^xi_j + Seta_ij G^eta_j
// with G = adj(J) F_K(U_j), the nodal flux as the stage kernel forms it.
template <int P, int K, int ns>
__device__ __forceinline__ void small_volume(const double* S, int e, double a00, double a01, double a10,
                                             double a11, ctab_t T, double (&r)[(P + 1) * (P + 2) / 2]) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    constexpr double gmi = 1.4 - 1.0;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const double rho = S[(0 * NP + j) * ns + e];
        const double m1 = S[(1 * NP + j) * ns + e];
        const double m2 = S[(2 * NP + j) * ns + e];
        const double m3 = S[(3 * NP + j) * ns + e];
        const double ir = fast_rcp(rho);
        const double uu = m1 * ir, vv = m2 * ir, E = m3 * ir;
        const double p = rho * gmi * (E - 0.5 * (uu * uu + vv * vv));
        double fx, fy;
        if constexpr (K == 0) { fx = rho * uu; fy = rho * vv; }
        else if constexpr (K == 1) { fx = rho * uu * uu + p; fy = rho * uu * vv; }
        else if constexpr (K == 2) { fx = rho * uu * vv; fy = rho * vv * vv + p; }
        else { const double H = E + p * ir; fx = rho * uu * H; fy = rho * vv * H; }
        const double gx = a00 * fx + a01 * fy;
        const double gy = a10 * fx + a11 * fy;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            r[i] = fma(T[tab_s(P) + (j * NP + i) * 2], gx, r[i]);
            r[i] = fma(T[tab_s(P) + (j * NP + i) * 2 + 1], gy, r[i]);
        }
    }
}

// a * b rounded to double on its own: the asm hides the product from the
// compiler, so it cannot be fused into a following add.
__device__ __forceinline__ double rounded_product(double a, double b) {
    double p = a * b;
    asm volatile("" : "+v"(p));
    return p;
}

template <int P, int NQ, int NEP>
__global__ void __launch_bounds__(kSmallThreads, 1) small_run_kernel(SmallParams prm) {
    constexpr int NT = kSmallThreads;
    constexpr int NP = (P + 1) * (P + 2) / 2;
    constexpr int NE = P + 1;
    extern __shared__ double lds[];
    // nep = ne rounded up to 64 (NEP, a launch parameter), so that every LDS
    // row offset is a compile-time constant
    constexpr int nep = NEP, ns = NEP + 1;
    const SmallLayout L = small_layout(NP, NE, NQ, NEP);
    // first padded phase-A item of each side group
    int ib[4];
    ib[0] = 0;
#pragma unroll
    for (int g = 0; g < 3; ++g) ib[g + 1] = ib[g] + ((prm.fbeg[g + 1] - prm.fbeg[g]) * NQ + 63) / 64 * 64;
    double* S = lds + L.S;
    double* Fs = lds + L.Fs;
    double* Nf = lds + L.Nf;
    double* Tw = lds + L.Tw;
    int* Fc = reinterpret_cast<int*>(lds + L.Fc);
    unsigned long long* key = reinterpret_cast<unsigned long long*>(lds + L.key);
    const int tid = static_cast<int>(threadIdx.x);
    const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
    const ctab_t T = const_table(prm.tab);

    // Phase B item of this thread: variable kb of element eb (a wavefront
    // covers 64 elements of one variable, so kb is wavefront-uniform).
    const int kb = __builtin_amdgcn_readfirstlane(tid / nep), eb = tid - kb * nep;
    const bool bact = tid < kSmallItems && kb < 4 && eb < prm.ne;
    double un[NP], sk[NP];  // U_n and k1 + 2 k2 + 2 k3 of (kb, eb)
    double a00 = 0, a01 = 0, a10 = 0, a11 = 0, idet = 0, len[3] = {0, 0, 0};
    if (tid == 0) *key = kNoError;
    for (int i = tid; i < 4 * NP * nep; i += NT) {
        const int v = i / nep, e = i - v * nep;
        S[v * ns + e] = e < prm.ne ? plane_at(prm.Un, v, ld8, static_cast<unsigned>(e) * 8u) : 0.0;
    }
    for (int i = tid; i < 3 * NQ * NE; i += NT) {
        Tw[i] = prm.tab[tab_tp(P) + i];
        Tw[3 * NQ * NE + i] = prm.tab[tab_tm(P, NQ) + i];
    }
    for (int f = tid; f < prm.nfaces; f += NT) {
        const int fc = prm.faces[f];
        const int e = fc & 1023, s = fc >> 20, s1 = s == 2 ? 0 : s + 1;
        const unsigned oe = static_cast<unsigned>(e) * 8u;
        const double dx = plane_at(prm.vtx, 2 * s1, ld8, oe) - plane_at(prm.vtx, 2 * s, ld8, oe);
        const double dy = plane_at(prm.vtx, 2 * s1 + 1, ld8, oe) - plane_at(prm.vtx, 2 * s + 1, ld8, oe);
        const double il = fast_rsq(dx * dx + dy * dy);
        Fc[f] = fc;
        Nf[f] = dy * il;
        Nf[L.nfp + f] = -dx * il;
    }
    if (bact) {
        const unsigned oe = static_cast<unsigned>(eb) * 8u;
        const double x0 = plane_at(prm.vtx, 0, ld8, oe), y0 = plane_at(prm.vtx, 1, ld8, oe);
        const double x1 = plane_at(prm.vtx, 2, ld8, oe), y1 = plane_at(prm.vtx, 3, ld8, oe);
        const double x2 = plane_at(prm.vtx, 4, ld8, oe), y2 = plane_at(prm.vtx, 5, ld8, oe);
        a00 = y2 - y0; a01 = x0 - x2; a10 = y0 - y1; a11 = x1 - x0;
        idet = 1.0 / ((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0));
        const double vx[3] = {x0, x1, x2}, vy[3] = {y0, y1, y2};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const double dx = vx[(s + 1) % 3] - vx[s], dy = vy[(s + 1) % 3] - vy[s];
            const double l2s = dx * dx + dy * dy;
            len[s] = l2s * fast_rsq(l2s);
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            un[j] = plane_at(prm.Un, kb * NP + j, ld8, oe);
            sk[j] = 0.0;
        }
    }
    __syncthreads();

    const int nstages = prm.method == 1 ? 4 * prm.nsteps : prm.nsteps;
    if (__builtin_amdgcn_readfirstlane(tid) >= kSmallItems) {
        // Role A: per stage one flux evaluation per listed side and Gauss
        // point, then the same two barriers as role B and the same stop after
        // the first failing stage.  The sides are listed by side index s,
        // each group's items padded to whole wavefronts, so s is uniform over
        // a wavefront; one item per thread and round.
        constexpr int NA = kSmallThreads - kSmallItems;
        const int w0 = __builtin_amdgcn_readfirstlane(tid - kSmallItems) & ~63;
        for (int st = 0; st < nstages; ++st) {
            for (int it0 = w0; it0 < ib[3]; it0 += NA) {
                const int it = it0 + (tid & 63);
                const int sd = __builtin_amdgcn_readfirstlane((it0 >= ib[1]) + (it0 >= ib[2]));
                const int li = it - ib[sd];
                if (li >= (prm.fbeg[sd + 1] - prm.fbeg[sd]) * NQ) continue;
                const int f = prm.fbeg[sd] + li / NQ, q = li - (li / NQ) * NQ;
                if (sd == 0) small_side_flux<P, NQ, 0, ns>(S, Fs, Nf, Tw, Fc, key, L.nfp, f, q, st);
                else if (sd == 1) small_side_flux<P, NQ, 1, ns>(S, Fs, Nf, Tw, Fc, key, L.nfp, f, q, st);
                else small_side_flux<P, NQ, 2, ns>(S, Fs, Nf, Tw, Fc, key, L.nfp, f, q, st);
            }
            __syncthreads();
            if (*key != kNoError) break;
            __syncthreads();
        }
    } else {
        // Role B: the volume term beside role A's fluxes, then phase B.
        for (int st = 0; st < nstages; ++st) {
            const int sub = prm.method == 1 ? (st & 3) : 0;  // RK4 stage within the step
            // The volume term reads only the stage state: formed before the
            // barrier, beside role A's flux evaluations.
            double r[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) r[i] = 0.0;
            if (bact) {
                switch (kb) {
                case 0: small_volume<P, 0, ns>(S, eb, a00, a01, a10, a11, T, r); break;
                case 1: small_volume<P, 1, ns>(S, eb, a00, a01, a10, a11, T, r); break;
                case 2: small_volume<P, 2, ns>(S, eb, a00, a01, a10, a11, T, r); break;
                default: small_volume<P, 3, ns>(S, eb, a00, a01, a10, a11, T, r); break;
                }
            }
            __syncthreads();
            if (*key != kNoError) break;  // the reference stops at its first failure

            // Phase B: the lift of the stored fluxes, Minv_ref / detJ, the update.
            double unew[NP];
            if (bact) {
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    const ctab_t tl = T + tab_lift(P, NQ) + s * NQ * NE;
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        // len * F once per point, then one FMA per edge node (the
                        // reference's len * w * phi * F regrouped at the ulp level).
                        const double lF = len[s] * Fs[((s * NQ + q) * 4 + kb) * ns + eb];
#pragma unroll
                        for (int m = 0; m < NE; ++m) r[edge_node<P>(s, m)] -= tl[q * NE + m] * lF;
                    }
                }
                const double dt = prm.dt;
#pragma unroll
                for (int i = 0; i < NP; ++i) {
                    double d = 0.0;
#pragma unroll
                    for (int j = 0; j < NP; ++j) d += T[tab_minv(P) + i * NP + j] * r[j];
                    const double kk = dt * (d * idet);
                    // The reference's update expressions, each operation rounded
                    // on its own (Timestepping.cpp:137,179,221,279, forward_euler
                    // :50); k/2 and 2k are exact, so only U_n + (1/6)(...) needs
                    // its multiply kept apart from its add.
                    if (prm.method == 0) {
                        unew[i] = S[(kb * NP + i) * ns + eb] + kk;
                    } else if (sub == 0) {
                        sk[i] = kk;
                        unew[i] = un[i] + kk / 2.0;
                    } else if (sub == 1) {
                        sk[i] = sk[i] + 2.0 * kk;
                        unew[i] = un[i] + kk / 2.0;
                    } else if (sub == 2) {
                        sk[i] = sk[i] + 2.0 * kk;
                        unew[i] = un[i] + kk;
                    } else {
                        un[i] = un[i] + rounded_product(1.0 / 6.0, sk[i] + kk);
                        unew[i] = un[i];
                    }
                }
            }
            // No barrier before the write: every read of another element's state
            // (phase A's traces, the volume term) ended at the barrier above, and
            // phase B reads only this thread's own entries of S.  The barrier after
            // it orders the write before the next stage's reads, and this stage's
            // reads of Fs before the next stage's phase A overwrites it.
            if (bact) {
#pragma unroll
                for (int i = 0; i < NP; ++i) S[(kb * NP + i) * ns + eb] = unew[i];
            }
            __syncthreads();
        }
    }

    // The state at the end of the launch (or of the failing stage) back to U_n.
    for (int i = tid; i < 4 * NP * nep; i += NT) {
        const int v = i / nep, e = i - v * nep;
        if (e < prm.ne) plane_at(prm.Un, v, ld8, static_cast<unsigned>(e) * 8u) = S[v * ns + e];
    }
    if (tid == 0 && *key != kNoError) atomicMin(prm.err, *key);
}

template <int NQ, int P, int NEP>
hipError_t launch_small_pn(const SmallParams& prm, hipStream_t st) {
    const size_t lds = sizeof(double) * small_layout((P + 1) * (P + 2) / 2, P + 1, NQ, NEP).total;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&small_run_kernel<P, NQ, NEP>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((small_run_kernel<P, NQ, NEP>), dim3(1), dim3(kSmallThreads), lds, st, prm);
    return hipGetLastError();
}

template <int NQ, int P>
hipError_t launch_small_p(const SmallParams& prm, hipStream_t st) {
    if (prm.ne <= 64) return launch_small_pn<NQ, P, 64>(prm, st);
    if (prm.ne <= 128) return launch_small_pn<NQ, P, 128>(prm, st);
    return hipErrorInvalidValue;
}

template <int NQ>
hipError_t NqLaunch<NQ>::small(const SmallParams& prm, int p, hipStream_t st) {
    switch (p) {
    case 0: return launch_small_p<NQ, 0>(prm, st);
    case 1: return launch_small_p<NQ, 1>(prm, st);
    case 2: return launch_small_p<NQ, 2>(prm, st);
    case 3: return launch_small_p<NQ, 3>(prm, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace vt
