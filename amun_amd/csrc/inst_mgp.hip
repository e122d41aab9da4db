// inst_mgp.hip -- kernel tables of the multi-dimensional Gaussian-process
// reconstruction for every equation system, Riemann solver, dimension and
// precision (compiled on its own so that the instantiation units build in
// parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_mgp(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_MGP>(d, c); }
}  // namespace amun_dev
