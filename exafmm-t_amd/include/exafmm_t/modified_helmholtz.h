// Modified Helmholtz (Yukawa) kernel FMM, e^{-kr}/(4 pi r)
// (reference include/modified_helmholtz.h:11-175).  Operators are level
// dependent (reference include/fmm.h).  The default constructor sets
// wavek = 0 (the reference leaves it uninitialised, modified_helmholtz.h:15).
#pragma once
#include "build_list.h"
#include "build_tree.h"
#include "fmm_base.h"

namespace exafmm_t {

class ModifiedHelmholtzFmm : public FmmBase<real_t> {
 public:
  real_t wavek = 0;
  ModifiedHelmholtzFmm() {}
  ModifiedHelmholtzFmm(int p_, int ncrit_, real_t wavek_, std::string filename_ = std::string())
      : FmmBase<real_t>(p_, ncrit_, filename_), wavek(wavek_) {
    if (filename.empty()) filename = "modified_helmholtz_d_p" + std::to_string(p) + ".dat";
  }
  int kernel_kind() const override { return XFMM_MODIFIED_HELMHOLTZ; }
  complex_t kernel_wavek() const override { return complex_t(wavek, 0); }
};

}  // namespace exafmm_t
