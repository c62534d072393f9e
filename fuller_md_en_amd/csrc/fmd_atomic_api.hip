// fmd_atomic_api.hip -- host side of the atomic (AIREBO) engine.
#include "fmd_engine.hpp"

namespace fmd {

int engine_prepare_atomic(fmd_engine *) {
  return set_error(FMD_ERR_STATE, "atomic engine: not available");
}
int engine_run_atomic(fmd_engine *, int, int, const fmd_step_ctrl *, fmd_thermo *) {
  return set_error(FMD_ERR_STATE, "atomic engine: not available");
}
int engine_forces_atomic(fmd_engine *, double *, double *, double *, double *) {
  return set_error(FMD_ERR_STATE, "atomic engine: not available");
}
void engine_free_atomic(fmd_engine *) {}

}  // namespace fmd
