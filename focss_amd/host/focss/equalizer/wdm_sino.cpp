#include "focss/equalizer/wdm_sino.h"

#include <algorithm>
#include <cassert>
#include <iostream>
#include <stdexcept>

#include "focss/runtime.h"

namespace focss {
namespace {
// rows 2..3 (the channel of interest) of a 6-envelope field
Field coi(const Field& f) {
    if (f.modes() != 6) throw std::invalid_argument("WdmPPE: fields hold 6 envelope rows");
    Field out(Domain{2, f.samples(), f.time_step(), f.center_frequency()});
    for (int p = 0; p < 2; ++p)
        for (int s = 0; s < f.samples(); ++s) out(p, s) = f(2 + p, s);
    return out;
}
}  // namespace

WdmPPE::WdmPPE() : c_radius_(0), d_radius_(0), trained_(false), intercept_(0), linear_weight_(0) {}
WdmPPE::WdmPPE(const int& radii) : WdmPPE() {
    c_radius_ = radii;
    d_radius_ = radii;
}
WdmPPE::WdmPPE(const int& c_radius, const int& d_radius) : WdmPPE() {
    c_radius_ = c_radius;
    d_radius_ = d_radius;
}
WdmPPE& WdmPPE::set_c_radius(const int& c_radius) {
    c_radius_ = c_radius;
    trained_ = false;
    return *this;
}
WdmPPE& WdmPPE::set_d_radius(const int& d_radius) {
    d_radius_ = d_radius;
    trained_ = false;
    return *this;
}

// wdm_sino.cpp:29-141 with the feature / Gram work on the device
void WdmPPE::train(const Field& tx, const Field& rx) {
    if (tx.modes() != 6 || rx.modes() != 6 || tx.samples() != rx.samples())
        throw std::invalid_argument("WdmPPE::train: tx and rx must be 6 envelope rows of equal length");
    const size_t bytes = (size_t)tx.size() * sizeof(complex_t);
    DeviceBuffer dtx(tx.raw(), bytes), drx(rx.raw(), bytes);
    const int cs = 2 * std::max(c_radius_, 0) + 1, ds = 2 * std::max(d_radius_, 0) + 1;
    C_.assign(c_radius_ >= 0 ? (size_t)cs * cs : 0, complex_t(0, 0));
    D_.assign(d_radius_ >= 0 ? (size_t)ds * ds : 0, complex_t(0, 0));
    double ic[2], lw[2];
    check(focss_wdm_ppe_train(runtime(), c_radius_, d_radius_, tx.samples(), dtx.ptr, drx.ptr,
                              reinterpret_cast<double*>(C_.data()), reinterpret_cast<double*>(D_.data()), ic, lw),
          "focss_wdm_ppe_train");
    intercept_ = complex_t(ic[0], ic[1]);
    linear_weight_ = complex_t(lw[0], lw[1]);
    trained_ = true;
    std::cout << "> wdm-ppe has been trained (" << c_radius_ << ',' << d_radius_ << ')' << std::endl;
}

complex_t WdmPPE::get_intercept() const {
    assert(trained_);
    return intercept_;
}
complex_t WdmPPE::get_linear_weight() const {
    assert(trained_);
    return linear_weight_;
}
complex_t WdmPPE::get_coo_weight() const {
    assert(trained_ && c_radius_ >= 0);
    return C_[(size_t)c_radius_ * (2 * c_radius_ + 1) + c_radius_];
}

void WdmPPE::terms(const Field& f, Field* spm, Field* xpm) const {
    assert(trained_);
    if (f.modes() != 6) throw std::invalid_argument("WdmPPE: fields hold 6 envelope rows");
    const size_t out_bytes = 2 * (size_t)f.samples() * sizeof(complex_t);
    DeviceBuffer din(f.raw(), (size_t)f.size() * sizeof(complex_t)), ds(out_bytes), dx(out_bytes);
    check(focss_wdm_ppe(runtime(), FOCSS_F64, c_radius_, reinterpret_cast<const double*>(C_.data()), d_radius_,
                        reinterpret_cast<const double*>(D_.data()), f.samples(), din.ptr, ds.ptr, dx.ptr),
          "focss_wdm_ppe");
    const Domain d2{2, f.samples(), f.time_step(), f.center_frequency()};
    if (spm) {
        *spm = Field(d2);
        ds.download(spm->raw());
    }
    if (xpm) {
        *xpm = Field(d2);
        dx.download(xpm->raw());
    }
}

// wdm_sino.cpp:246-300
Field WdmPPE::linear(const Field& f) const {
    assert(trained_);
    return linear_weight_ * coi(f);
}
Field WdmPPE::spm(const Field& f) const {
    Field s;
    terms(f, &s, nullptr);
    return s;
}
Field WdmPPE::xpm(const Field& f) const {
    Field x;
    terms(f, nullptr, &x);
    return x;
}
Field WdmPPE::cubic(const Field& f) const {
    Field s, x;
    terms(f, &s, &x);
    return s + x;
}
Field WdmPPE::model(const Field& f) const { return linear(f) + cubic(f); }
Field WdmPPE::equalize(const Field& tx, const Field& rx) const { return (coi(rx) - cubic(tx)) / linear_weight_; }
Field WdmPPE::equalize(const Field& tx_spm, const Field& tx_xpm, const Field& rx) const {
    return (coi(rx) - spm(tx_spm) - xpm(tx_xpm)) / linear_weight_;
}
}  // namespace focss
