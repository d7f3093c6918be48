#include "focss/module/fiber.h"

namespace focss {
// derived fibres                                    fiber.cpp:4-45
Fiber Fiber::linear() const {
    Fiber f = *this;
    f.gamma = 0;
    return f;
}

Fiber Fiber::nonlinear() const {
    Fiber f = *this;
    f.alpha = 0;
    f.beta2 = 0;
    return f;
}

Fiber Fiber::dispersive() const {
    Fiber f = *this;
    f.alpha = 0;
    f.gamma = 0;
    return f;
}

Fiber Fiber::operator-() const {
    Fiber f = *this;
    f.alpha = -alpha;
    f.beta2 = -beta2;
    f.gamma = -gamma;
    return f;
}

// a negative factor reverses the fibre and scales by |factor|
Fiber Fiber::operator*(const double& k) const {
    Fiber f = k >= 0 ? *this : -*this;
    f.length = length * (k < 0 ? -k : k);
    return f;
}
}  // namespace focss
