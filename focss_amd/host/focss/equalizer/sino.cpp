#include "focss/equalizer/sino.h"

#include <cassert>
#include <cmath>
#include <fstream>
#include <iostream>

#include "focss/runtime.h"

namespace focss {
SINO::SINO() : radius_(0), lambda_(1), type_(ORDINARY), trained_(false), linear_weight_(0) {}
SINO::SINO(const int& r) : SINO() { radius_ = r; }
SINO::SINO(const int& r, const double& l) : SINO() {
    radius_ = r;
    lambda_ = l;
}
SINO::SINO(const int& r, const double& l, const SinoType& t) : SINO() {
    radius_ = r;
    lambda_ = l;
    type_ = t;
}
SINO& SINO::set_radius(const int& r) {
    radius_ = r;
    trained_ = false;
    return *this;
}
SINO& SINO::set_lambda(const double& l) {
    lambda_ = l;
    trained_ = false;
    return *this;
}
SINO& SINO::set_type(const SinoType& t) {
    type_ = t;
    trained_ = false;
    return *this;
}
SINO& SINO::set_linear_weight(const complex_t& w) {
    linear_weight_ = w;
    trained_ = false;
    return *this;
}


// sino.cpp:42-159 with the feature / Gram work on the device
void SINO::train(const Field& tx, const Field& rx) {
    assert(tx.domain() == rx.domain());
    DeviceBuffer dtx(tx.raw(), (size_t)tx.size() * sizeof(complex_t));
    DeviceBuffer drx(rx.raw(), (size_t)rx.size() * sizeof(complex_t));
    int side = 2 * radius_ + 1;
    weights_.assign((size_t)side * side, complex_t(0, 0));
    double lw_in[2] = {linear_weight_.real(), linear_weight_.imag()};
    double ic[2], lw[2];
    check(focss_sino_train(runtime(), radius_, (int)type_, lw_in, tx.modes(), tx.samples(), dtx.ptr, drx.ptr,
                           reinterpret_cast<double*>(weights_.data()), ic, lw),
          "focss_sino_train");
    intercept_ = complex_t(ic[0], ic[1]);
    linear_weight_ = complex_t(lw[0], lw[1]);
    trained_ = true;
    std::cout << "> sino has been trained (" << radius_ << ")" << std::endl;
}

complex_t SINO::get_intercept() const { return intercept_; }
complex_t SINO::get_linear_weight() const { return linear_weight_; }
complex_t SINO::get_coo_weight() const { return weights_[(size_t)radius_ * (2 * radius_ + 1) + radius_]; }

Field SINO::model(const Field& f) const { return linear(f) + cubic(f); }
Field SINO::equalize(const Field& rx, const Field& tx) const { return (rx - cubic(tx)) / linear_weight_; }
Field SINO::linear(const Field& f) const { return linear_weight_ * f; }

// sino.cpp:203-220 on the device
Field SINO::cubic(const Field& f) const {
    assert(trained_);
    DeviceBuffer din(f.raw(), (size_t)f.size() * sizeof(complex_t)), dout((size_t)f.size() * sizeof(complex_t));
    check(focss_sino_cubic(runtime(), FOCSS_F64, radius_, reinterpret_cast<const double*>(weights_.data()), f.modes(),
                           f.samples(), din.ptr, dout.ptr),
          "focss_sino_cubic");
    Field out(f.domain());
    dout.download(out.raw());
    return out;
}

void SINO::save_matrix(const std::string& filename) {
    std::ofstream out(filename);
    save_matrix(out);
}

// real, imaginary, modulus and argument blocks       sino.cpp:229-268
void SINO::save_matrix(std::ostream& out) {
    int side = 2 * radius_ + 1;
    for (int block = 0; block < 4; ++block) {
        for (int m = 0; m < side; ++m) {
            for (int n = 0; n < side; ++n) {
                complex_t w = weights_[(size_t)m * side + n];
                double v = block == 0 ? w.real() : block == 1 ? w.imag() : block == 2 ? std::abs(w) : std::arg(w);
                out << v << ' ';
            }
            out << '\n';
        }
        if (block < 3) out << "\n\n";
    }
}
}  // namespace focss
