// focss::SINO -- single-channel / dual-polarisation perturbation-triplet
// equalizer with the reference's interface (src/focss/equalizer/sino.h:8-50).
// Training (features, Gram matrix) and the cubic term run on the GPU through
// focss_sino_train / focss_sino_cubic; weights are kept as a dense
// (2R+1)^2 matrix W(m, n) at [(R+m)(2R+1) + R+n].
#ifndef FOCSS_EQUALIZER_SINO_H_
#define FOCSS_EQUALIZER_SINO_H_

#include <ostream>
#include <string>
#include <vector>

#include "focss/field.h"

namespace focss {
enum SinoType { ORDINARY, FIXED, HOMOGENEOUS };

class SINO {
  public:
    SINO();
    SINO(const int& radius);
    SINO(const int& radius, const double& lambda);
    SINO(const int& radius, const double& lambda, const SinoType& type);
    SINO& set_radius(const int& radius);
    SINO& set_lambda(const double& lambda);
    SINO& set_type(const SinoType& type);
    SINO& set_linear_weight(const complex_t& linear_weight);

    void train(const Field& transmitted, const Field& received);

    complex_t get_intercept() const;
    complex_t get_linear_weight() const;
    complex_t get_coo_weight() const;
    Field model(const Field& original) const;
    Field equalize(const Field& rx, const Field& tx) const;
    Field linear(const Field& original) const;
    Field cubic(const Field& original) const;

    void save_matrix(const std::string& filename);
    void save_matrix(std::ostream& output);

    const std::vector<complex_t>& weights() const { return weights_; }

  private:
    int radius_;
    double lambda_;  // kept for interface parity; the reference's solve ignores it
    SinoType type_;
    bool trained_;
    complex_t intercept_;
    complex_t linear_weight_;
    std::vector<complex_t> weights_;
};
}  // namespace focss

#endif  // FOCSS_EQUALIZER_SINO_H_
