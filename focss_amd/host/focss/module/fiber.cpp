#include "focss/module/fiber.h"

#include <initializer_list>

namespace focss {
namespace {
// the fibre with the attenuation, dispersion or nonlinearity switched off
// where the flag is false (fiber.cpp:4-22)
Fiber keeping(Fiber f, bool loss, bool dispersion, bool kerr) {
    if (!loss) f.alpha = 0;
    if (!dispersion) f.beta2 = 0;
    if (!kerr) f.gamma = 0;
    return f;
}
}  // namespace

Fiber Fiber::linear() const { return keeping(*this, true, true, false); }
Fiber Fiber::nonlinear() const { return keeping(*this, false, false, true); }
Fiber Fiber::dispersive() const { return keeping(*this, false, true, false); }

// the backward fibre: every propagation coefficient negated (fiber.cpp:24-30)
Fiber Fiber::operator-() const {
    Fiber f = *this;
    for (double* c : {&f.alpha, &f.beta2, &f.gamma}) *c = -*c;
    return f;
}

// |k| times the length; a k that is not >= 0 also reverses the fibre
// (fiber.cpp:32-45)
Fiber Fiber::operator*(const double& k) const {
    const bool forward = k >= 0;
    Fiber f = forward ? *this : -*this;
    f.length = length * (forward ? k : -k);
    return f;
}
}  // namespace focss
