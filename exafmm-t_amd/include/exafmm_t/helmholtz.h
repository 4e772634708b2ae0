// Low-frequency Helmholtz kernel FMM, e^{ikr}/(4 pi r) with complex k and
// complex charges (reference include/helmholtz.h:11-211).  Operators are
// level dependent (reference include/fmm.h).
#pragma once
#include "build_list.h"
#include "build_tree.h"
#include "fmm_base.h"

namespace exafmm_t {

class HelmholtzFmm : public FmmBase<complex_t> {
 public:
  complex_t wavek = complex_t(0, 0);
  HelmholtzFmm() {}
  HelmholtzFmm(int p_, int ncrit_, complex_t wavek_, std::string filename_ = std::string())
      : FmmBase<complex_t>(p_, ncrit_, filename_), wavek(wavek_) {
    if (filename.empty()) filename = "helmholtz_d_p" + std::to_string(p) + ".dat";
  }
  int kernel_kind() const override { return XFMM_HELMHOLTZ; }
  complex_t kernel_wavek() const override { return wavek; }
};

}  // namespace exafmm_t
