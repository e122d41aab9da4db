// inst_unit.hip -- one kernel-table unit: the sweep, primitive, integral,
// unit-test, multi-level and adjoint kernels of one (reconstruction,
// equation system, precision) entry of inst_list.h.  The build compiles this
// file once per entry, with -DAMUN_INST_ENTRY=<entry name> and the entry's
// values as -DAMUN_INST_REC, -DAMUN_INST_EQ and -DAMUN_INST_PREC, so the
// units build in parallel.
#include "kernels.h"

#if !defined(AMUN_INST_ENTRY) || !defined(AMUN_INST_REC) || !defined(AMUN_INST_EQ) || \
    !defined(AMUN_INST_PREC)
#error "compile with the -DAMUN_INST_* values of one entry of inst_list.h"
#endif

namespace amun_dev {

bool AMUN_INST_ENTRY(Dispatch& d, const amun_cfg& c) {
  using T = std::conditional_t<(AMUN_INST_PREC) == AMUN_PREC_MIXED, float, double>;
  return pick_eq<(AMUN_INST_REC), (AMUN_INST_EQ), T>(d, c);
}

}  // namespace amun_dev
