// Small fixed-size float vectors and matrices used by the host-side API types.
// They provide the member access the gs_native API uses on its Eigen types
// (operator[], x()/y()/z(), operator()(row, column), norm(), Zero()).
#pragma once

#include <array>
#include <cmath>

namespace gs_native {

template<int Size>
struct VectorF {
    std::array<float, Size> v{};

    static VectorF Zero() {
        return VectorF{};
    }
    float& operator[](int i) {
        return v[i];
    }
    float operator[](int i) const {
        return v[i];
    }
    float& x() {
        return v[0];
    }
    float& y() {
        return v[1];
    }
    float& z() {
        return v[2];
    }
    float x() const {
        return v[0];
    }
    float y() const {
        return v[1];
    }
    float z() const {
        return v[2];
    }
    float squaredNorm() const {
        float s = 0.0f;
        for (float e : v) {
            s += e * e;
        }
        return s;
    }
    float norm() const {
        return std::sqrt(squaredNorm());
    }
    bool allFinite() const {
        for (float e : v) {
            if (!std::isfinite(e)) {
                return false;
            }
        }
        return true;
    }
};

using Vector3f = VectorF<3>;
using Vector4f = VectorF<4>;

inline Vector3f MakeVector3f(float x, float y, float z) {
    Vector3f out;
    out.v = {x, y, z};
    return out;
}

template<int Rows, int Cols>
struct MatrixF {
    std::array<float, Rows * Cols> m{};

    static MatrixF Zero() {
        return MatrixF{};
    }
    static MatrixF Identity() {
        MatrixF out;
        for (int i = 0; i < (Rows < Cols ? Rows : Cols); ++i) {
            out(i, i) = 1.0f;
        }
        return out;
    }
    float& operator()(int r, int c) {
        return m[r * Cols + c];
    }
    float operator()(int r, int c) const {
        return m[r * Cols + c];
    }
    bool allFinite() const {
        for (float e : m) {
            if (!std::isfinite(e)) {
                return false;
            }
        }
        return true;
    }
};

using Matrix3f = MatrixF<3, 3>;

} // namespace gs_native
