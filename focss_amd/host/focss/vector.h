// Owning-or-viewing contiguous vector with elementwise arithmetic: the
// container type of the focss C++ API (src/focss/vector.h:8-282).  A view
// made by Vector::proxy() aliases foreign storage, e.g. one mode of a Field.
#ifndef FOCSS_VECTOR_H_
#define FOCSS_VECTOR_H_

#include <cassert>
#include <initializer_list>
#include <vector>

namespace focss {
template <typename Scalar>
class Vector {
  public:
    Vector() : view_(nullptr), view_size_(0) {}
    explicit Vector(const int& size) : store_(size), view_(nullptr), view_size_(0) {}
    Vector(const int& size, const Scalar& fill) : store_(size, fill), view_(nullptr), view_size_(0) {}
    Vector(std::initializer_list<Scalar> list) : store_(list), view_(nullptr), view_size_(0) {}
    Vector(const Vector& o) : store_(o.begin_(), o.begin_() + o.size()), view_(nullptr), view_size_(0) {}
    Vector(Vector&& o) : view_(nullptr), view_size_(0) {
        if (o.view_) {
            store_.assign(o.view_, o.view_ + o.view_size_);
        } else {
            store_.swap(o.store_);
        }
    }
    Vector& operator=(const Vector& o) {
        if (this == &o) return *this;
        if (view_) {
            assert(view_size_ == o.size());
            for (int i = 0; i < view_size_; ++i) view_[i] = o[i];
        } else {
            store_.assign(o.begin_(), o.begin_() + o.size());
        }
        return *this;
    }
    Vector& operator=(Vector&& o) {
        if (view_ || o.view_) return *this = static_cast<const Vector&>(o);
        store_.swap(o.store_);
        return *this;
    }
    virtual ~Vector() {}

    static Vector proxy(const int& size, Scalar* data) {
        Vector v;
        v.view_ = data;
        v.view_size_ = size;
        return v;
    }
    bool is_owner() const { return view_ == nullptr; }

    int size() const { return view_ ? view_size_ : (int)store_.size(); }
    Scalar* raw() { return view_ ? view_ : store_.data(); }
    const Scalar* raw() const { return view_ ? view_ : store_.data(); }
    Scalar& operator[](const int& i) { return raw()[i]; }
    // by value, as the reference's const subscript (vector.h:169): a caller
    // may then call non-const members of the copy
    Scalar operator[](const int& i) const { return raw()[i]; }

    template <typename T>
    Vector& operator*=(const T& s) {
        for (int i = 0; i < size(); ++i) raw()[i] *= s;
        return *this;
    }
    template <typename T>
    Vector& operator/=(const T& s) {
        for (int i = 0; i < size(); ++i) raw()[i] /= s;
        return *this;
    }
    Vector& operator+=(const Vector& o) {
        assert(size() == o.size());
        for (int i = 0; i < size(); ++i) raw()[i] += o[i];
        return *this;
    }
    Vector& operator-=(const Vector& o) {
        assert(size() == o.size());
        for (int i = 0; i < size(); ++i) raw()[i] -= o[i];
        return *this;
    }

  private:
    const Scalar* begin_() const { return raw(); }
    std::vector<Scalar> store_;
    Scalar* view_;
    int view_size_;
};
}  // namespace focss

#endif  // FOCSS_VECTOR_H_
