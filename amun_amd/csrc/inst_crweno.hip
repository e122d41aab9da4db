// inst_crweno.hip -- kernel tables of the compact reconstructions
// (CRWENO5-Z, -YC, -NS; CRMP5, CRMP5LD, CRMP7) for every equation system, Riemann solver, dimension
// and precision (compiled on its own so that the instantiation units build
// in parallel).
#include "kernels.h"

namespace amun_dev {
bool pick_crweno(Dispatch& d, const amun_cfg& c) {
  switch (c.reconstruction) {
  case AMUN_REC_CRWENO5Z: return pick_rec<REC_CRWENO5Z>(d, c);
  case AMUN_REC_CRWENO5YC: return pick_rec<REC_CRWENO5YC>(d, c);
  case AMUN_REC_CRMP5: return pick_rec<REC_CRMP5>(d, c);
  case AMUN_REC_CRMP5LD: return pick_rec<REC_CRMP5LD>(d, c);
  case AMUN_REC_CRMP7: return pick_rec<REC_CRMP7>(d, c);
  default: return pick_rec<REC_CRWENO5NS>(d, c);
  }
}
}  // namespace amun_dev
