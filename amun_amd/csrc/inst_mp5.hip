// inst_mp5.hip -- kernel table of the MP5 reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_mp5(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_MP5>(d, c); }
}  // namespace amun_dev
