// inst_weno3.hip -- kernel table of the WENO3 reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_weno3(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_WENO3>(d, c); }
}  // namespace amun_dev
