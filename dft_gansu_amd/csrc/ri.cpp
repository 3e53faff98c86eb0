// RI (density fitting) path of the Fock engine.
#include "engine.hpp"

namespace gansu {

struct FockEngine::RIState {};
void ri_destroy(FockEngine::RIState* p) { delete p; }

void FockEngine::precompute_ri() { throw Error(GANSU_E_UNSUPPORTED, "RI path not available in this build"); }

void FockEngine::build_ri(const double*, double*, double*, const double*, double*, hipStream_t) {
  throw Error(GANSU_E_UNSUPPORTED, "RI path not available in this build");
}

}  // namespace gansu
