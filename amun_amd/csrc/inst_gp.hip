// inst_gp.hip -- kernel table of the GP reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_gp(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_GP>(d, c); }
}  // namespace amun_dev
