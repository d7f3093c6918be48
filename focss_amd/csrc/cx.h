// Complex arithmetic for the HIP kernels: interleaved (re, im) pairs, the
// same storage as std::complex<double> / std::complex<float>, so a field is
// the reference's mode-major complex buffer (src/focss/field.cpp:140-155).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace focss_hip {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
    R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cx<R> mk(R a, R b) {
    cx<R> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) {
    return mk<R>(a.x + b.x, a.y + b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) {
    return mk<R>(a.x - b.x, a.y - b.y);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
    return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, R s) {
    return mk<R>(a.x * s, a.y * s);
}
template <typename R>
__host__ __device__ __forceinline__ cx<R> conj(cx<R> a) {
    return mk<R>(a.x, -a.y);
}
template <typename R>
__host__ __device__ __forceinline__ R norm2(cx<R> a) {
    return a.x * a.x + a.y * a.y;
}
// multiply by (sign * i): forward transforms use sign = -1
template <typename R, int SIGN>
__host__ __device__ __forceinline__ cx<R> mul_si(cx<R> a) {
    return SIGN > 0 ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

}  // namespace focss_hip
