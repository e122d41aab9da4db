// inst_ppm.hip -- kernel table of the PPM reconstruction for every equation
// system, Riemann solver, dimension and precision (compiled on its own so
// that the instantiation units build in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_ppm(Dispatch& d, const amun_cfg& c) { return pick_rec<REC_PPM>(d, c); }
}  // namespace amun_dev
