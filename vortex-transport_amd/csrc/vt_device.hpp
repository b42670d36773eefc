// vt_device.hpp -- layout shared by the host set-up and the HIP kernels.
#pragma once

namespace vt {

// Packed operator table for one (P, NQ), all doubles, laid out so that each
// kernel loop reads one contiguous run per iteration (scalar loads):
//   S[NP][NP][2]      S[j][i] = {Sxi[i][j], Seta[i][j]}   stiffness, column-major pairs
//   Minv[NP][NP]      Minv_ref, row-major                 (forward: rows)
//   MinvT[NP][NP]     Minv_ref transposed                 (adjoint: columns)
//   tp[3][NQ][P+1]    own-side trace weights  phi_{edge node m}(plus point q)
//   tm[3][NQ][P+1]    neighbour trace weights phi_{edge node m}(minus point q)
//   lift[3][NQ][P+1]  tp * w_q
constexpr int tab_np(int p) { return (p + 1) * (p + 2) / 2; }
constexpr int tab_s(int) { return 0; }
constexpr int tab_minv(int p) { return 2 * tab_np(p) * tab_np(p); }
constexpr int tab_minvt(int p) { return 3 * tab_np(p) * tab_np(p); }
constexpr int tab_tp(int p) { return 4 * tab_np(p) * tab_np(p); }
constexpr int tab_tm(int p, int nq) { return tab_tp(p) + 3 * nq * (p + 1); }
constexpr int tab_lift(int p, int nq) { return tab_tm(p, nq) + 3 * nq * (p + 1); }
constexpr int tab_size(int p, int nq) { return tab_lift(p, nq) + 3 * nq * (p + 1); }

// Edge nodes per side (side 1 = v0->v1 at eta = 0, side 2 = v1->v2 on
// xi + eta = 1, side 3 = v2->v0 at xi = 0), nodes numbered eta-outer/xi-inner.
constexpr int kEdge0[3][1] = {{0}, {0}, {0}};
constexpr int kEdge1[3][2] = {{0, 1}, {1, 2}, {2, 0}};
constexpr int kEdge2[3][3] = {{0, 1, 2}, {2, 4, 5}, {5, 3, 0}};
constexpr int kEdge3[3][4] = {{0, 1, 2, 3}, {3, 6, 8, 9}, {9, 7, 4, 0}};

// Error key written by the kernels with atomicMin: the smallest key is the
// first failure in the reference's serial order (element, then quadrature
// point, then side; Evolve.cpp:215-221), within the earliest stage.
//   key = stage << 40 | (((elem * 16 + q) * 4 + side) * 4 + code)
// code: 1 left state, 2 right state, 3 Roe c2 (Numericalflux.cpp:51,84,120).
constexpr unsigned long long kNoError = ~0ull;

}  // namespace vt
