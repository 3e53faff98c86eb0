// Instances of the two-electron backward kernel for part of the quartet
// classes (split over four translation units so that they compile in parallel).
#include "gradient_2e_impl.hpp"

namespace gansu {

bool grad_2e_dispatch_3(int key, const ClassLaunch& c, const Grad2eArgs& a, int maxnp, int maxnk,
                          hipStream_t s, hipError_t* err) {
  using namespace grad_detail;
  switch (key) {
    GANSU_GRAD_CASE(1, 1, 0, 0)
    GANSU_GRAD_CASE(2, 0, 1, 0)
    GANSU_GRAD_CASE(2, 1, 1, 0)
    GANSU_GRAD_CASE(2, 1, 2, 1)
    GANSU_GRAD_CASE(2, 2, 2, 1)
    default:
      return false;
  }
}

}  // namespace gansu
