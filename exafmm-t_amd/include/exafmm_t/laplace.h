// Laplace kernel FMM, 1/(4 pi r) (reference include/laplace.h:11-167).
#pragma once
#include "build_list.h"
#include "build_tree.h"
#include "fmm_base.h"

namespace exafmm_t {

class LaplaceFmm : public FmmBase<real_t> {
 public:
  LaplaceFmm() {}
  LaplaceFmm(int p_, int ncrit_, std::string filename_ = std::string()) : FmmBase<real_t>(p_, ncrit_, filename_) {
    if (filename.empty()) filename = "laplace_d_p" + std::to_string(p) + ".dat";
  }
  int kernel_kind() const override { return XFMM_LAPLACE; }
};

}  // namespace exafmm_t
