// Small complex FP64 helpers shared by the channel (exciton) kernels: capacitance
// determinants and inverses of at most 3 x 3.
#pragma once
#include <hip/hip_runtime.h>

namespace gz {

template <typename T> struct cpx;
template <> struct cpx<float> { using t = float2; };
template <> struct cpx<double> { using t = double2; };

struct zc {  // FP64 complex for the small determinants
    double x, y;
};
__device__ __forceinline__ zc zmk(double x, double y) { zc r; r.x = x; r.y = y; return r; }
__device__ __forceinline__ zc zadd(zc a, zc b) { return zmk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ zc zsub(zc a, zc b) { return zmk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ zc zmul(zc a, zc b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// conj(a) * b
__device__ __forceinline__ zc zconjmul(zc a, zc b) { return zmk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ double znorm(zc a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ zc zdiv(zc a, zc b)
{
    const double n = znorm(b);
    return zmk((a.x * b.x + a.y * b.y) / n, (a.y * b.x - a.x * b.y) / n);
}
template <typename C> __device__ __forceinline__ zc toz(C v) { return zmk((double)v.x, (double)v.y); }

// v_l + v_(l xor 16) and v_l + v_(l xor 32) in every lane (every lane active), by the gfx950
// lane swaps: v_permlane16_swap(v, v) returns {v_l, v_(l xor 16)} in some order, and IEEE
// addition is commutative, so the sums equal those of a ds_bpermute butterfly
__device__ __forceinline__ float sum_xor16(float v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float sum_xor32(float v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ double sum_xor16(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0])) +
           __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ double sum_xor32(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    return __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0])) +
           __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}

// one DPP move of a 64-bit value (two 32-bit halves); lanes of rows outside ROW_MASK keep old
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v, unsigned long long old)
{
    const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)old, (int)(unsigned)v, CTRL, ROW_MASK, 0xF, false);
    const int hi =
        __builtin_amdgcn_update_dpp((int)(unsigned)(old >> 32), (int)(unsigned)(v >> 32), CTRL, ROW_MASK, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// one step of the wave reduction below: sum s; largest v with the smallest index k on ties
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ void sum_argmax_step(double &s, double &v, int &k)
{
    s += __longlong_as_double((long long)dpp_u64<CTRL, ROW_MASK>((unsigned long long)__double_as_longlong(s), 0ull));
    const double v2 = __longlong_as_double((long long)dpp_u64<CTRL, ROW_MASK>(
        (unsigned long long)__double_as_longlong(v), (unsigned long long)__double_as_longlong(v)));
    const int k2 = __builtin_amdgcn_update_dpp(k, k, CTRL, ROW_MASK, 0xF, false);
    if (v2 > v || (v2 == v && k2 < k)) {
        v = v2;
        k = k2;
    }
}

// the sum and the argmax over the 64 lanes of a wavefront, left in lane 63 (every lane
// active): quad permutations, half-row and row mirrors, then the row broadcasts 15 and 31
__device__ __forceinline__ void wave_sum_argmax_to_lane63(double &s, double &v, int &k)
{
    sum_argmax_step<0xB1>(s, v, k);       // quad_perm [1, 0, 3, 2]
    sum_argmax_step<0x4E>(s, v, k);       // quad_perm [2, 3, 0, 1]
    sum_argmax_step<0x141>(s, v, k);      // row_half_mirror
    sum_argmax_step<0x140>(s, v, k);      // row_mirror: every lane holds its row's result
    sum_argmax_step<0x142, 0xA>(s, v, k); // row_bcast:15 into rows 1 and 3
    sum_argmax_step<0x143, 0xC>(s, v, k); // row_bcast:31 into rows 2 and 3
}

// running product of complex factors as mantissa (mr, mi) times 2^ex: the mantissa is
// rescaled by a power of two only when its exponent leaves [-256, 256], so that a walk needs
// no logarithm per accepted move
__device__ __forceinline__ void zamp_mul(double &mr, double &mi, double &ex, double qr, double qi)
{
    const double re = mr * qr - mi * qi, im = mr * qi + mi * qr;
    const int be = (int)((__double_as_longlong(fmax(fabs(re), fabs(im))) >> 52) & 0x7FF) - 1023;
    if (be > 256 || be < -256) {
        const double sc = __longlong_as_double((long long)(1023 - be) << 52);  // 2^-be
        mr = re * sc;
        mi = im * sc;
        ex += be;
    } else {
        mr = re;
        mi = im;
    }
}

__device__ inline zc det_small(int n, const zc *B)  // n <= 3, row-major n x n
{
    if (n == 0) return zmk(1.0, 0.0);
    if (n == 1) return B[0];
    if (n == 2) return zsub(zmul(B[0], B[3]), zmul(B[1], B[2]));
    zc t0 = zsub(zmul(B[4], B[8]), zmul(B[5], B[7]));
    zc t1 = zsub(zmul(B[3], B[8]), zmul(B[5], B[6]));
    zc t2 = zsub(zmul(B[3], B[7]), zmul(B[4], B[6]));
    return zadd(zsub(zmul(B[0], t0), zmul(B[1], t1)), zmul(B[2], t2));
}

// inverse of an n x n (n <= 3) complex matrix by the adjugate
__device__ inline void inv_small(int n, const zc *B, zc *Bi)
{
    const zc d = det_small(n, B);
    if (n == 1) { Bi[0] = zdiv(zmk(1.0, 0.0), d); return; }
    if (n == 2) {
        Bi[0] = zdiv(B[3], d);
        Bi[1] = zdiv(zmk(-B[1].x, -B[1].y), d);
        Bi[2] = zdiv(zmk(-B[2].x, -B[2].y), d);
        Bi[3] = zdiv(B[0], d);
        return;
    }
    auto m = [&](int r, int c) { return B[r * 3 + c]; };
    auto cof = [&](int r0, int r1, int c0, int c1) {
        return zsub(zmul(m(r0, c0), m(r1, c1)), zmul(m(r0, c1), m(r1, c0)));
    };
    // Bi[i][j] = C[j][i] / d
    zc C[9];
    C[0] = cof(1, 2, 1, 2);
    C[1] = zsub(zmk(0, 0), cof(1, 2, 0, 2));
    C[2] = cof(1, 2, 0, 1);
    C[3] = zsub(zmk(0, 0), cof(0, 2, 1, 2));
    C[4] = cof(0, 2, 0, 2);
    C[5] = zsub(zmk(0, 0), cof(0, 2, 0, 1));
    C[6] = cof(0, 1, 1, 2);
    C[7] = zsub(zmk(0, 0), cof(0, 1, 0, 2));
    C[8] = cof(0, 1, 0, 1);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Bi[i * 3 + j] = zdiv(C[j * 3 + i], d);
}

// 3 x 3 determinant and inverse (a smaller capacitance is padded with the identity, which
// leaves its determinant and its inverse block unchanged)
__device__ __forceinline__ zc det3(const zc *B)
{
    const zc t0 = zsub(zmul(B[4], B[8]), zmul(B[5], B[7]));
    const zc t1 = zsub(zmul(B[3], B[8]), zmul(B[5], B[6]));
    const zc t2 = zsub(zmul(B[3], B[7]), zmul(B[4], B[6]));
    return zadd(zsub(zmul(B[0], t0), zmul(B[1], t1)), zmul(B[2], t2));
}
__device__ __forceinline__ void inv3(const zc *B, zc *Bi)
{
    const zc d = det3(B);
    // one reciprocal of the determinant, nine products (a division per cofactor would put
    // eighteen FP64 divides on the one lane that runs this)
    const double rn = 1.0 / (d.x * d.x + d.y * d.y);
    const zc di = zmk(d.x * rn, -d.y * rn);
    auto cof = [&](int r0, int r1, int c0, int c1) {
        return zsub(zmul(B[r0 * 3 + c0], B[r1 * 3 + c1]), zmul(B[r0 * 3 + c1], B[r1 * 3 + c0]));
    };
    // Bi[i][j] = cofactor C[j][i] / d
    Bi[0] = zmul(cof(1, 2, 1, 2), di);
    Bi[3] = zmul(zsub(zmk(0, 0), cof(1, 2, 0, 2)), di);
    Bi[6] = zmul(cof(1, 2, 0, 1), di);
    Bi[1] = zmul(zsub(zmk(0, 0), cof(0, 2, 1, 2)), di);
    Bi[4] = zmul(cof(0, 2, 0, 2), di);
    Bi[7] = zmul(zsub(zmk(0, 0), cof(0, 2, 0, 1)), di);
    Bi[2] = zmul(cof(0, 1, 1, 2), di);
    Bi[5] = zmul(zsub(zmk(0, 0), cof(0, 1, 0, 2)), di);
    Bi[8] = zmul(cof(0, 1, 0, 1), di);
}
__device__ __forceinline__ void eye3(zc *B)
{
#pragma unroll
    for (int i = 0; i < 9; ++i) B[i] = zmk((i % 4) == 0 ? 1.0 : 0.0, 0.0);
}

}  // namespace gz
