// inst_tvd.hip -- kernel table of the TVD reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_tvd(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_TVD>(d, c); }
}  // namespace amun_dev
