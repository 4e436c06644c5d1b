#pragma once
// Small fixed-size vector / matrix types for the host-side cloth API.
//
// The reference headers (include/cloth_mesh.h:11-12, sim_constraints.h:25,
// mesh_topology.h:7) expose Eigen types in their CPU members. Eigen is not a
// dependency here, so these types provide the subset of that interface the
// cloth API and its callers use: element access through operator()(i) and
// operator()(r, c) (column-major storage, like Eigen's default), x()/y()/z(),
// dot, norm, squaredNorm, normalized, 2x2 inverse and determinant.

#include <cmath>
#include <type_traits>

namespace cloth {

template <typename S, int N>
struct Vec {
    S v[N];

    Vec() { for (int i = 0; i < N; ++i) v[i] = S(0); }
    template <int M = N, typename = typename std::enable_if<M == 3>::type>
    Vec(S a, S b, S c) { v[0] = a; v[1] = b; v[2] = c; }
    template <int M = N, typename = typename std::enable_if<M == 4>::type>
    Vec(S a, S b, S c, S d) { v[0] = a; v[1] = b; v[2] = c; v[3] = d; }

    static Vec Zero() { return Vec(); }

    S&       operator()(int i)       { return v[i]; }
    const S& operator()(int i) const { return v[i]; }
    S&       operator[](int i)       { return v[i]; }
    const S& operator[](int i) const { return v[i]; }
    S x() const { return v[0]; }
    S y() const { return v[1]; }
    S z() const { return v[2]; }

    Vec operator+(const Vec& o) const { Vec r; for (int i = 0; i < N; ++i) r.v[i] = v[i] + o.v[i]; return r; }
    Vec operator-(const Vec& o) const { Vec r; for (int i = 0; i < N; ++i) r.v[i] = v[i] - o.v[i]; return r; }
    Vec operator-() const { Vec r; for (int i = 0; i < N; ++i) r.v[i] = -v[i]; return r; }
    Vec operator*(S s) const { Vec r; for (int i = 0; i < N; ++i) r.v[i] = v[i] * s; return r; }
    Vec operator/(S s) const { Vec r; for (int i = 0; i < N; ++i) r.v[i] = v[i] / s; return r; }
    Vec& operator+=(const Vec& o) { for (int i = 0; i < N; ++i) v[i] += o.v[i]; return *this; }
    Vec& operator-=(const Vec& o) { for (int i = 0; i < N; ++i) v[i] -= o.v[i]; return *this; }
    bool operator==(const Vec& o) const { for (int i = 0; i < N; ++i) if (v[i] != o.v[i]) return false; return true; }
    bool operator!=(const Vec& o) const { return !(*this == o); }

    S dot(const Vec& o) const { S s = v[0] * o.v[0]; for (int i = 1; i < N; ++i) s += v[i] * o.v[i]; return s; }
    S squaredNorm() const { return dot(*this); }
    S norm() const { return std::sqrt(squaredNorm()); }
    // Divides each component by the norm; a zero vector is returned unchanged.
    Vec normalized() const {
        const S n2 = squaredNorm();
        if (n2 > S(0)) return *this / std::sqrt(n2);
        return *this;
    }
};

template <typename S, int N>
inline Vec<S, N> operator*(S s, const Vec<S, N>& a) { return a * s; }

// Column-major R x C matrix: operator()(r, c) reads m[c * R + r].
template <typename S, int R, int C>
struct Mat {
    S m[R * C];

    Mat() { for (int i = 0; i < R * C; ++i) m[i] = S(0); }
    static Mat Zero() { return Mat(); }
    static Mat Identity() { Mat r; for (int i = 0; i < (R < C ? R : C); ++i) r(i, i) = S(1); return r; }

    S&       operator()(int r, int c)       { return m[c * R + r]; }
    const S& operator()(int r, int c) const { return m[c * R + r]; }

    Vec<S, R> col(int c) const { Vec<S, R> o; for (int r = 0; r < R; ++r) o.v[r] = (*this)(r, c); return o; }
    void set_col(int c, const Vec<S, R>& o) { for (int r = 0; r < R; ++r) (*this)(r, c) = o.v[r]; }

    Mat<S, C, R> transpose() const {
        Mat<S, C, R> t;
        for (int r = 0; r < R; ++r) for (int c = 0; c < C; ++c) t(c, r) = (*this)(r, c);
        return t;
    }
    template <int K>
    Mat<S, R, K> operator*(const Mat<S, C, K>& o) const {
        Mat<S, R, K> p;
        for (int r = 0; r < R; ++r)
            for (int k = 0; k < K; ++k) {
                S s = (*this)(r, 0) * o(0, k);
                for (int c = 1; c < C; ++c) s += (*this)(r, c) * o(c, k);
                p(r, k) = s;
            }
        return p;
    }
    Vec<S, R> operator*(const Vec<S, C>& x) const {
        Vec<S, R> y;
        for (int r = 0; r < R; ++r) {
            S s = (*this)(r, 0) * x.v[0];
            for (int c = 1; c < C; ++c) s += (*this)(r, c) * x.v[c];
            y.v[r] = s;
        }
        return y;
    }
    Mat operator-(const Mat& o) const { Mat r; for (int i = 0; i < R * C; ++i) r.m[i] = m[i] - o.m[i]; return r; }
    Mat operator+(const Mat& o) const { Mat r; for (int i = 0; i < R * C; ++i) r.m[i] = m[i] + o.m[i]; return r; }
    Mat operator*(S s) const { Mat r; for (int i = 0; i < R * C; ++i) r.m[i] = m[i] * s; return r; }
    // Frobenius norm.
    S norm() const { S s = S(0); for (int i = 0; i < R * C; ++i) s += m[i] * m[i]; return std::sqrt(s); }

    // 2x2 only.
    S determinant() const {
        static_assert(R == 2 && C == 2, "determinant: 2x2 only");
        return (*this)(0, 0) * (*this)(1, 1) - (*this)(1, 0) * (*this)(0, 1);
    }
    Mat inverse() const {
        static_assert(R == 2 && C == 2, "inverse: 2x2 only");
        const S inv_det = S(1) / determinant();
        Mat r;
        r(0, 0) =  (*this)(1, 1) * inv_det;
        r(1, 0) = -(*this)(1, 0) * inv_det;
        r(0, 1) = -(*this)(0, 1) * inv_det;
        r(1, 1) =  (*this)(0, 0) * inv_det;
        return r;
    }
};

using Vec3f   = Vec<float, 3>;
using Vec3i   = Vec<int, 3>;
using Vec4i   = Vec<int, 4>;
using Mat2f   = Mat<float, 2, 2>;
using Mat3x2f = Mat<float, 3, 2>;
using Mat6x9f = Mat<float, 6, 9>;

}  // namespace cloth
