// Small in-register DFT butterflies (radix 2, 3, 4, 5, 8), forward and inverse.
#pragma once
#include "common.h"

namespace fans {

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cplx<T> (&v)[R]);


template <>
__device__ __forceinline__ void dft<double, 2, false>(cplx<double> (&v)[2])
{
    cplx<double> a = v[0];
    v[0]           = a + v[1];
    v[1]           = a - v[1];
}
template <>
__device__ __forceinline__ void dft<double, 2, true>(cplx<double> (&v)[2])
{
    dft<double, 2, false>(v);
}
template <>
__device__ __forceinline__ void dft<float, 2, false>(cplx<float> (&v)[2])
{
    cplx<float> a = v[0];
    v[0]          = a + v[1];
    v[1]          = a - v[1];
}
template <>
__device__ __forceinline__ void dft<float, 2, true>(cplx<float> (&v)[2])
{
    dft<float, 2, false>(v);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cplx<T> (&v)[3])
{
    const T      s3 = (T) 0.86602540378443864676;
    cplx<T>      t1 = v[1] + v[2];
    cplx<T>      t2 = v[0] - cscale(t1, (T) 0.5);
    cplx<T>      t3 = mul_mi<T, INV>(cscale(v[1] - v[2], s3));
    v[0]            = v[0] + t1;
    v[1]            = t2 + t3;
    v[2]            = t2 - t3;
}
template <>
__device__ __forceinline__ void dft<double, 3, false>(cplx<double> (&v)[3]) { dft3<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 3, true>(cplx<double> (&v)[3]) { dft3<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 3, false>(cplx<float> (&v)[3]) { dft3<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 3, true>(cplx<float> (&v)[3]) { dft3<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx<T> (&v)[4])
{
    cplx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
    cplx<T> t3 = mul_mi<T, INV>(v[1] - v[3]);
    v[0]       = t0 + t2;
    v[2]       = t0 - t2;
    v[1]       = t1 + t3;
    v[3]       = t1 - t3;
}
template <>
__device__ __forceinline__ void dft<double, 4, false>(cplx<double> (&v)[4]) { dft4<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 4, true>(cplx<double> (&v)[4]) { dft4<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 4, false>(cplx<float> (&v)[4]) { dft4<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 4, true>(cplx<float> (&v)[4]) { dft4<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft5(cplx<T> (&v)[5])
{
    const T c1 = (T) 0.30901699437494742410, c2 = (T) -0.80901699437494742410;
    const T s1 = (T) 0.95105651629515357212, s2 = (T) 0.58778525229247312917;
    cplx<T> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
    cplx<T> p1 = v[0] + cscale(a1, c1) + cscale(a2, c2);
    cplx<T> p2 = v[0] + cscale(a1, c2) + cscale(a2, c1);
    cplx<T> q1 = mul_mi<T, INV>(cscale(b1, s1) + cscale(b2, s2));
    cplx<T> q2 = mul_mi<T, INV>(cscale(b1, s2) - cscale(b2, s1));
    v[0]       = v[0] + a1 + a2;
    v[1]       = p1 + q1;
    v[4]       = p1 - q1;
    v[2]       = p2 + q2;
    v[3]       = p2 - q2;
}
template <>
__device__ __forceinline__ void dft<double, 5, false>(cplx<double> (&v)[5]) { dft5<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 5, true>(cplx<double> (&v)[5]) { dft5<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 5, false>(cplx<float> (&v)[5]) { dft5<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 5, true>(cplx<float> (&v)[5]) { dft5<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cplx<T> (&v)[8])
{
    const T r = (T) 0.70710678118654752440;
    cplx<T> a[4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        a[t] = v[t] + v[t + 4];
        b[t] = v[t] - v[t + 4];
    }
    // b[1] *= W8, b[2] *= W8^2, b[3] *= W8^3  (W8 = exp(-+ 2 pi i / 8))
    {
        cplx<T> x = b[1];
        b[1]      = INV ? cplx<T>{(x.x - x.y) * r, (x.x + x.y) * r} : cplx<T>{(x.x + x.y) * r, (x.y - x.x) * r};
        b[2]      = mul_mi<T, INV>(b[2]);
        x         = b[3];
        b[3]      = INV ? cplx<T>{(-x.x - x.y) * r, (x.x - x.y) * r} : cplx<T>{(x.y - x.x) * r, (-x.x - x.y) * r};
    }
    dft4<T, INV>(a);
    dft4<T, INV>(b);
    v[0] = a[0];
    v[2] = a[1];
    v[4] = a[2];
    v[6] = a[3];
    v[1] = b[0];
    v[3] = b[1];
    v[5] = b[2];
    v[7] = b[3];
}
template <>
__device__ __forceinline__ void dft<double, 8, false>(cplx<double> (&v)[8]) { dft8<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 8, true>(cplx<double> (&v)[8]) { dft8<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 8, false>(cplx<float> (&v)[8]) { dft8<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 8, true>(cplx<float> (&v)[8]) { dft8<float, true>(v); }

}  // namespace fans
