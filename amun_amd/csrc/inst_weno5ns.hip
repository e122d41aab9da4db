// inst_weno5ns.hip -- kernel table of the WENO5-NS reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_weno5ns(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_WENO5NS>(d, c); }
}  // namespace amun_dev
