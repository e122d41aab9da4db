// inst_weno5yc.hip -- kernel table of the WENO5-YC reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_weno5yc(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_WENO5YC>(d, c); }
}  // namespace amun_dev
