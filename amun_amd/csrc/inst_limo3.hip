// inst_limo3.hip -- kernel table of the LimO3 reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_limo3(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_LIMO3>(d, c); }
}  // namespace amun_dev
