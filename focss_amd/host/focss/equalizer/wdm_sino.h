// focss::WdmPPE -- the dual-polarisation WDM perturbation-triplet equalizer
// with the reference's interface (src/focss/equalizer/wdm_sino.h:8-52), on
// focss::Field in place of arma::cx_mat: the inputs hold six envelope rows
// (modes) -- channel * 2 + polarisation, rows 2..3 the channel of interest,
// the rows of load_field's matrix -- and the outputs two (the COI's x and
// y).  Training (features, Gram matrix) and spm / xpm run on the GPU through
// focss_wdm_ppe_train / focss_wdm_ppe.  C and D are dense (2c+1)^2 and
// (2d+1)^2 matrices, W(m, n) at [(r+m)(2r+1) + r+n]; a radius of -1 drops
// that term.
#ifndef FOCSS_EQUALIZER_WDM_SINO_H_
#define FOCSS_EQUALIZER_WDM_SINO_H_

#include <vector>

#include "focss/field.h"

namespace focss {
class WdmPPE {
  public:
    WdmPPE();
    WdmPPE(const int& radii);
    WdmPPE(const int& c_radius, const int& d_radius);
    WdmPPE& set_c_radius(const int& c_radius);
    WdmPPE& set_d_radius(const int& d_radius);

    void train(const Field& tx, const Field& rx);

    complex_t get_intercept() const;
    complex_t get_linear_weight() const;
    complex_t get_coo_weight() const;
    Field linear(const Field& field) const;
    Field spm(const Field& field) const;
    Field xpm(const Field& field) const;
    Field cubic(const Field& field) const;
    Field model(const Field& field) const;
    Field equalize(const Field& tx, const Field& rx) const;
    Field equalize(const Field& tx_spm, const Field& tx_xpm, const Field& rx) const;

    const std::vector<complex_t>& c_weights() const { return C_; }
    const std::vector<complex_t>& d_weights() const { return D_; }

  private:
    // spm and xpm of one 6-envelope field in one device call
    void terms(const Field& field, Field* spm, Field* xpm) const;

    int c_radius_;
    int d_radius_;
    bool trained_;
    complex_t intercept_;
    complex_t linear_weight_;
    std::vector<complex_t> C_;
    std::vector<complex_t> D_;
};
}  // namespace focss

#endif  // FOCSS_EQUALIZER_WDM_SINO_H_
