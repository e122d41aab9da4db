// inst_mp7.hip -- kernel table of the MP7 reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_mp7(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_MP7>(d, c); }
}  // namespace amun_dev
