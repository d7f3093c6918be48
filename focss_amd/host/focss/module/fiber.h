// focss::Fiber -- POD fibre in SI units (src/focss/module/fiber.h:5-21).
#ifndef FOCSS_MODULE_FIBER_H_
#define FOCSS_MODULE_FIBER_H_

namespace focss {
struct Fiber {
    double alpha;       // power attenuation [1/m]
    double beta2;       // [s^2/m]
    double gamma;       // [1/(W m)]
    double length;      // [m]
    double wavelength;  // [m]
    double kappa;       // polarisation coupling

    Fiber linear() const;
    Fiber nonlinear() const;
    Fiber dispersive() const;
    Fiber operator-() const;
    Fiber operator*(const double& length_factor) const;
};
}  // namespace focss

#endif  // FOCSS_MODULE_FIBER_H_
