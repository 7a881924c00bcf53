// Device-side math for the SPH passes: float3 helpers, the smoothing kernels,
// symmetric 3x3 storage and the fixed-sweep Jacobi eigen solver.
//
// Every expression keeps the operand order of the reference kernels
// (kernels/common.cl:78-350) and the library is compiled without FMA
// contraction, so each result is the IEEE-rounded value of the same sequence
// of operations the CPU oracle performs.
//
// Symmetric 3x3 matrices use the order [00, 11, 22, 01, 02, 12]
// (kernels/common.cl:27-35); full 3x3 matrices are row-major.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

namespace sphk {

struct V3 {
    float x, y, z;
};

// 12-byte tuple with 4-byte alignment: one global_load_dwordx3 per vector.
struct __attribute__((packed, aligned(4))) Packed3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
// Per-slot fields stay below 4 GiB (checked when a handle or a batch is
// created), so a slot's byte offset is a 32-bit value and a load is
// "scalar base + 32-bit vector offset", with no 64-bit address arithmetic.
template <typename T>
__device__ __forceinline__ const T *at_offset(const T *p, unsigned byte_offset) {
    return reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + byte_offset);
}
__device__ __forceinline__ V3 load3(const float *p, unsigned i) {
    const Packed3 t = *reinterpret_cast<const Packed3 *>(at_offset(p, i * 12u));
    return V3{t.x, t.y, t.z};
}
// The 12-byte tuple at a byte offset (the scans step these by 12 per slot).
__device__ __forceinline__ V3 load3_at(const float *p, unsigned byte_offset) {
    const Packed3 t = *reinterpret_cast<const Packed3 *>(at_offset(p, byte_offset));
    return V3{t.x, t.y, t.z};
}
// Element i of a per-slot array of 32-bit values.
template <typename T>
__device__ __forceinline__ T load1(const T *p, unsigned i) {
    return *at_offset(p, i * 4u);
}
__device__ __forceinline__ void store3(float *p, unsigned i, V3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
// Float division through a double reciprocal: float(double(a) * (1.0 / double(b)))
// is the correctly rounded a / b for every pair of floats (zeros, infinities,
// NaNs and subnormals included).  The exact quotient of two 24-bit
// significands is never a float midpoint and lies at least 2^-49 (relative)
// from every midpoint, in the normal and the subnormal range; the double
// product is within 2^-52 of the quotient, so rounding it to float cannot
// cross a midpoint.  Several quotients by one divisor then cost one
// reciprocal and a multiply each instead of one float division each.
//
// The reciprocal is 1.0 / double(b) correctly rounded, formed without the
// division sequence: the hardware estimate (relative error below 2^-24)
// refined by two Newton steps y' = y + y (1 - b y) in fused multiply-adds.
// For a float divisor this equals the correctly rounded quotient bit for bit:
// sph_hip_check_recip compares the two for all 2^32 bit patterns on the
// device.  Zeros, infinities and NaN make a Newton step NaN, and for them
// the estimate itself is already exact (+-inf, +-0, NaN).
__device__ __forceinline__ double recip(float b) {
    const double bd = static_cast<double>(b);
    const double y0 = __builtin_amdgcn_rcp(bd);
    const double y1 = __builtin_fma(y0, __builtin_fma(-bd, y0, 1.0), y0);
    const double y2 = __builtin_fma(y1, __builtin_fma(-bd, y1, 1.0), y1);
    return __builtin_isfinite(y2) ? y2 : y0;
}
// The same value through the compiler's division, for the self-check.
__device__ __forceinline__ double recip_by_division(float b) { return 1.0 / static_cast<double>(b); }
__device__ __forceinline__ float div_recip(float a, double rb) {
    return static_cast<float>(static_cast<double>(a) * rb);
}
__device__ __forceinline__ V3 div_recip(V3 a, double rb) {
    return V3{div_recip(a.x, rb), div_recip(a.y, rb), div_recip(a.z, rb)};
}
__device__ __forceinline__ V3 operator/(V3 a, float s) { return div_recip(a, recip(s)); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float sqlength(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(sqlength(a)); }

// Range prefilter.  sqlength_fused rounds twice instead of five times; with
// non-negative terms each of the two forms lies within a factor (1 +- u)^3 of
// the exact a.a (u = 2^-24), so sqlength(a) < r2 implies
// sqlength_fused(a) < r2 (1 + 6.1 u) <= widened_r2(r2).  NaN components make
// both forms NaN.  A scan that appends on the prefilter re-tests each hit
// with sqlength, so the set it keeps is the one sqlength selects.
__device__ __forceinline__ float sqlength_fused(V3 a) {
    return __builtin_fmaf(a.z, a.z, __builtin_fmaf(a.y, a.y, a.x * a.x));
}
__device__ __forceinline__ float widened_r2(float r2) { return r2 * (1.0f + 0x1p-20f); }

// Correctly rounded square root (the value of sqrtf) without the small-
// argument scaling of the compiler's sequence: the hardware estimate r, then
// the fused multiply-add residuals s - r'r of its one-ulp neighbours r' pick
// the rounded root.  That holds for s >= 2^-96 and for +-0 and +inf; other
// arguments (smaller positive s, negative s, NaN) take sqrtf, behind a
// wave-uniform branch that the pair passes practically never take.
__device__ __forceinline__ float sqrt_rn(float s) {
    const float r = __builtin_amdgcn_sqrtf(s);
    const float rm = __uint_as_float(__float_as_uint(r) - 1u);
    const float rp = __uint_as_float(__float_as_uint(r) + 1u);
    float out = __builtin_fmaf(-rm, r, s) <= 0.0f ? rm : r;
    out = __builtin_fmaf(-rp, r, s) > 0.0f ? rp : out;
    const bool other = !(s >= 0x1p-96f) && s != 0.0f;
    if (__builtin_expect(__any(other), 0)) {
        const float t = sqrtf(s);
        out = other ? t : out;
    }
    return out;
}

// Reflection of point `a` in the plane through `pt` with normal `nrm`:
// a - 2 n (n . (a - pt))   (kernels/common.cl:672).
__device__ __forceinline__ V3 reflect_point(V3 a, V3 pt, V3 nrm) {
    float d = dot(nrm, a - pt);
    return a - (2.0f * nrm) * d;
}
// Reflection of a direction: a - 2 n (n . a)   (kernels/solid_corotation.cl:358).
__device__ __forceinline__ V3 reflect_dir(V3 a, V3 nrm) {
    float d = dot(nrm, a);
    return a - (2.0f * nrm) * d;
}

// Smoothing kernels (Mueller et al. 2003) with host-precomputed constants
//   poly6 = 315/(64 pi h^9), spiky = -45/(pi h^6), lap = 45/(pi h^6).
struct KernelConsts {
    float h, h2, poly6, spiky, lap;
};

// W_poly6(r), zero beyond h   (kernels/common.cl:78-80)
__device__ __forceinline__ float w_poly6(float r, const KernelConsts &k) {
    if (r > k.h) return 0.0f;
    float t = k.h2 - r * r;
    return k.poly6 * (t * t * t);
}
// The same for r known to be <= h (a listed pair).
__device__ __forceinline__ float w_poly6_in_range(float r, const KernelConsts &k) {
    float t = k.h2 - r * r;
    return k.poly6 * (t * t * t);
}
// grad W_spiky(v)   (kernels/common.cl:81-87)
__device__ __forceinline__ V3 grad_spiky(V3 v, const KernelConsts &k) {
    float r = length(v);
    if (r > k.h) return V3{0.0f, 0.0f, 0.0f};
    V3 dir = v / r;
    float s = (k.h - r) * (k.h - r);
    return (dir * k.spiky) * s;
}
// grad_spiky(v) and grad_spiky(-v) together.  Negation is exact and every
// operation of grad_spiky is sign-symmetric under round-to-nearest, so
// grad_spiky(-v) is bitwise -grad_spiky(v) when |v| <= h (and +0 otherwise,
// as grad_spiky returns).
__device__ __forceinline__ void grad_spiky_pm(V3 v, const KernelConsts &k, V3 *plus, V3 *minus) {
    const float r = length(v);
    if (r > k.h) {
        *plus = *minus = V3{0.0f, 0.0f, 0.0f};
        return;
    }
    const V3 g = ((v / r) * k.spiky) * ((k.h - r) * (k.h - r));
    *plus = g;
    *minus = -g;
}
// lap W_visc(r)   (kernels/common.cl:88-90)
__device__ __forceinline__ float lap_visc(float r, const KernelConsts &k) {
    if (r > k.h) return 0.0f;
    return k.lap * (k.h - r);
}

// ---- 3x3 algebra ---------------------------------------------------------
enum { S00 = 0, S11 = 1, S22 = 2, S01 = 3, S02 = 4, S12 = 5 };

__device__ __forceinline__ int sym_index(int r, int c) {
    // [00,11,22,01,02,12]
    return r == c ? r : (r + c == 1 ? S01 : (r + c == 2 ? S02 : S12));
}

// s * v with s symmetric   (kernels/common.cl:230-238)
__device__ __forceinline__ V3 sym_mul(const float *s, V3 v) {
    return V3{s[S00] * v.x + s[S01] * v.y + s[S02] * v.z,
              s[S01] * v.x + s[S11] * v.y + s[S12] * v.z,
              s[S02] * v.x + s[S12] * v.y + s[S22] * v.z};
}
// m * v, m row-major   (kernels/common.cl:221-229)
__device__ __forceinline__ V3 mat_mul(const float *m, V3 v) {
    return V3{m[0] * v.x + m[1] * v.y + m[2] * v.z,
              m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// out = a * s, s symmetric   (kernels/common.cl:187-199)
template <typename T>
__device__ __forceinline__ void mul_full_sym(const T *a, const T *s, T *out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            out[3 * r + c] = a[3 * r] * s[sym_index(0, c)] + a[3 * r + 1] * s[sym_index(1, c)] +
                             a[3 * r + 2] * s[sym_index(2, c)];
}
// out = a * b, both full   (kernels/common.cl:208-220)
template <typename T>
__device__ __forceinline__ void mul_full(const T *a, const T *b, T *out) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            out[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
}
// P diag(d) P^T, symmetric result   (kernels/common.cl:169-186)
template <typename T>
__device__ __forceinline__ void p_d_pt(const T *d, const T *p, T *out) {
    T d0p00 = d[0] * p[0], d1p01 = d[1] * p[1], d2p02 = d[2] * p[2];
    T d0p10 = d[0] * p[3], d1p11 = d[1] * p[4], d2p12 = d[2] * p[5];
    T d0p20 = d[0] * p[6], d1p21 = d[1] * p[7], d2p22 = d[2] * p[8];
    out[S00] = p[0] * d0p00 + p[1] * d1p01 + p[2] * d2p02;
    out[S01] = p[0] * d0p10 + p[1] * d1p11 + p[2] * d2p12;
    out[S02] = p[0] * d0p20 + p[1] * d1p21 + p[2] * d2p22;
    out[S11] = p[3] * d0p10 + p[4] * d1p11 + p[5] * d2p12;
    out[S12] = p[3] * d0p20 + p[4] * d1p21 + p[5] * d2p22;
    out[S22] = p[6] * d0p20 + p[7] * d1p21 + p[8] * d2p22;
}
template <typename T>
__device__ __forceinline__ T det3(const T *m) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
           m[2] * (m[3] * m[7] - m[4] * m[6]);
}

__device__ __forceinline__ float t_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double t_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float t_abs(float x) { return fabsf(x); }
__device__ __forceinline__ double t_abs(double x) { return fabs(x); }

// The FP32 path's atan, sin and cos.  The reference calls its OpenCL
// platform's library (under -cl-fast-relaxed-math); here they are defined
// once, as double polynomials rounded to float, and the oracle writes out the
// same operation sequence, so the two agree bit for bit (within a float ulp
// of the C library):
//   atan: |x| > 1 -> pi/2 - atan(1/|x|); then t > tan(pi/12) ->
//         pi/6 + atan((t sqrt3 - 1) / (t + sqrt3)); then the Taylor series
//         t sum_{n<=15} (-t^2)^n / (2n+1) (|t| <= 0.268); odd in x;
//   sin, cos on |a| <= pi/4 + 1e-6 (the Jacobi half angle): Taylor series
//         to a^19 and a^20.
__device__ __forceinline__ double poly_atan(double x) {
    double t = fabs(x);
    const bool inv = t > 1.0;
    if (inv) t = 1.0 / t;
    const bool shift = t > 0.2679491924311227;
    if (shift) t = (t * 1.7320508075688772 - 1.0) / (t + 1.7320508075688772);
    const double u = t * t;
    double p = 1.0 / 31.0;
    p = -p;
#pragma unroll
    for (int n = 14; n >= 0; --n) p = p * u + ((n & 1) ? -1.0 : 1.0) / static_cast<double>(2 * n + 1);
    double r = t * p;
    if (shift) r = 0.5235987755982989 + r;
    if (inv) r = 1.5707963267948966 - r;
    return copysign(r, x);
}
__device__ __forceinline__ double poly_sin(double a) {
    const double c[10] = {1.0, -1.0 / 6.0, 1.0 / 120.0, -1.0 / 5040.0, 1.0 / 362880.0,
                          -1.0 / 39916800.0, 1.0 / 6227020800.0, -1.0 / 1307674368000.0,
                          1.0 / 355687428096000.0, -1.0 / 121645100408832000.0};
    const double u = a * a;
    double p = c[9];
#pragma unroll
    for (int n = 8; n >= 0; --n) p = p * u + c[n];
    return a * p;
}
__device__ __forceinline__ double poly_cos(double a) {
    const double c[11] = {1.0, -1.0 / 2.0, 1.0 / 24.0, -1.0 / 720.0, 1.0 / 40320.0,
                          -1.0 / 3628800.0, 1.0 / 479001600.0, -1.0 / 87178291200.0,
                          1.0 / 20922789888000.0, -1.0 / 6402373705728000.0,
                          1.0 / 2432902008176640000.0};
    const double u = a * a;
    double p = c[10];
#pragma unroll
    for (int n = 9; n >= 0; --n) p = p * u + c[n];
    return p;
}
// cos and sin of angle = 1/2 atan(e / d).  FP32: the reference's atan, cos and
// sin (poly_*) on the reference's quotient.  FP64 (the precision upgrade, not bound
// to the reference's bits): the same angle without library trig, from
// r = sqrt(d^2 + e^2): cos phi = |d| / r, sin phi = sign(d) e / r (phi in
// (-pi/2, pi/2), the range of atan; sign(-0) = -1 as e / -0 = -inf has it),
// cos(phi/2) = sqrt((1 + cos phi) / 2) >= cos(pi/4) and
// sin(phi/2) = sin phi / (2 cos(phi/2)).  d and e come from FP32 data, so
// d^2 + e^2 cannot overflow; a non-finite r (infinite or NaN input) takes
// the quotient form.  Agrees with the library calls to a few FP64 ulps.
__device__ __forceinline__ void half_atan_cs(float e, float d, float *c, float *s) {
    const float angle = 0.5f * static_cast<float>(poly_atan(static_cast<double>(e / d)));
    *c = static_cast<float>(poly_cos(static_cast<double>(angle)));
    *s = static_cast<float>(poly_sin(static_cast<double>(angle)));
}
__device__ __forceinline__ void half_atan_cs(double e, double d, double *c, double *s) {
    const double r = sqrt(d * d + e * e);
    double cphi, sphi;
    if (r <= 1.7e308) {
        const double inv = 1.0 / r;
        cphi = fabs(d) * inv;
        sphi = (signbit(d) ? -e : e) * inv;
    } else {
        const double x = e / d;
        if (fabs(x) <= 1.0) {
            const double q = 1.0 / sqrt(1.0 + x * x);
            cphi = q;
            sphi = x * q;
        } else {
            const double ix = 1.0 / fabs(x);
            const double q = 1.0 / sqrt(1.0 + ix * ix);
            cphi = ix * q;
            sphi = copysign(q, x);
        }
    }
    const double ch = sqrt(0.5 + 0.5 * cphi);
    *c = ch;
    *s = sphi / (ch + ch);
}

// One Givens rotation that zeroes off-diagonal `od` (3, 4 or 5) of the
// symmetric `a`, applied to the eigenvector columns of `p`
// (kernels/common.cl:95-155).  angle = 1/2 atan(2 a_xy / (a_xx - a_yy)); when
// a_xy is exactly 0 the rotation is the identity (the reference evaluates
// 0/0 there when a_xx == a_yy).
template <typename T>
__device__ __forceinline__ void givens(const T *a, const T *p, int od, T *ao, T *po) {
    const int k = od - 3;
    const int ixx = k == 2 ? 1 : 0;
    const int iyy = k == 0 ? 1 : 2;
    const int iee = 2 - k;
    const int iex = k == 0 ? 4 : 3;
    const int iey = k == 2 ? 4 : 5;
    T c = T(1), s = T(0);   // cos(0), sin(0)
    if (a[od] != T(0)) half_atan_cs(T(2) * a[od], a[ixx] - a[iyy], &c, &s);
    const T c2 = c * c, s2 = s * s, cs = c * s;
    ao[ixx] = a[ixx] * c2 + a[iyy] * s2 + T(2) * a[od] * cs;
    ao[od] = -a[ixx] * cs + a[iyy] * cs + a[od] * (c2 - s2);
    ao[iyy] = a[ixx] * s2 + a[iyy] * c2 - T(2) * a[od] * cs;
    ao[iee] = a[iee];
    ao[iex] = a[iex] * c + a[iey] * s;
    ao[iey] = -a[iex] * s + a[iey] * c;
    for (int r = 0; r < 3; ++r) {
        po[3 * r + ixx] = p[3 * r + ixx] * c + p[3 * r + iyy] * s;
        po[3 * r + iyy] = -p[3 * r + ixx] * s + p[3 * r + iyy] * c;
        po[3 * r + iee] = p[3 * r + iee];
    }
}

// Index (3..5) of the largest |off-diagonal|, first one on ties, 3 if all 0
// (kernels/common.cl:156-168).
template <typename T>
__device__ __forceinline__ int max_offdiag(const T *a) {
    T m = T(0);
    int im = 0;
    for (int q = 0; q < 3; ++q) {
        T v = t_abs(a[3 + q]);
        if (v > m) { m = v; im = q; }
    }
    return im + 3;
}

// Eight max-pivot rotations (kernels/common.cl:257-295 with iterations = 8,
// threshold = 0): eigenvalues in ev, eigenvectors in the columns of p.
template <typename T>
__device__ __forceinline__ void jacobi8(const T *a_in, T *ev, T *p) {
    T a[6], b[6], q[9];
    for (int t = 0; t < 6; ++t) a[t] = a_in[t];
    for (int t = 0; t < 9; ++t) p[t] = (t % 4 == 0) ? T(1) : T(0);
    int od = max_offdiag(a);
#pragma unroll
    for (int it = 0; it < 8; it += 2) {
        givens(a, p, od, b, q);
        od = max_offdiag(b);
        givens(b, q, od, a, p);
        od = max_offdiag(a);
    }
    ev[0] = a[0];
    ev[1] = a[1];
    ev[2] = a[2];
}

}  // namespace sphk
