// vt_kernels_nq.hip -- instantiates the stage kernels of one quadrature
// size, VT_NQ (compiled once per NQ = 1..10).
#include "vt_stage.hpp"
#include "vt_adjoint.hpp"
#include "vt_small.hpp"

namespace vt {
template struct NqLaunch<VT_NQ>;
}  // namespace vt
