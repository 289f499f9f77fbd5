// Device-side q-mesh arithmetic shared by the kernels.
//
// Mesh point index k = c2 + n2*(c1 + n1*c0), c in [0,n); reduced coordinate c/n - 1/2
// (THERMACOND's Gamma-containing even meshes, make_kp_reg / make_kp_reg_tet).  Sums and
// differences of mesh points are again mesh points, so every get_k_info of the reference
// (rest.f90:3024-3058) becomes integer arithmetic:
//   q1 + q2  -> c1 + c2 - n/2 (mod n), inside the first zone iff n/2 <= c1+c2 < 3n/2
//   q1 - q2  -> c1 - c2 + n/2 (mod n), inside iff -n/2 <= c1-c2 < n/2
//   -q1 - q2 -> 3n/2 - c1 - c2 (mod n);  Gamma at c = n/2
// Division by the runtime mesh sizes uses a float reciprocal: floor((k + 1/2) / n) is exact
// while the float product error (~1.2e-7 k/n) stays below 1/(2n), i.e. for k < 4e6.
#pragma once
#include "alt3_impl.h"

namespace alt3 {

struct MeshDev {
    int n0, n1, n2, n12;
    float inv2, inv12;
    int h0, h1, h2;
    static MeshDev make(const int *nc)
    {
        MeshDev m;
        m.n0 = nc[0];
        m.n1 = nc[1];
        m.n2 = nc[2];
        m.n12 = nc[1] * nc[2];
        m.inv2 = 1.0f / (float)nc[2];
        m.inv12 = 1.0f / (float)m.n12;
        m.h0 = nc[0] / 2;
        m.h1 = nc[1] / 2;
        m.h2 = nc[2] / 2;
        return m;
    }
    __device__ __forceinline__ void coords(int k, int &c0, int &c1, int &c2) const
    {
        c0 = (int)(((float)k + 0.5f) * inv12);
        int r = k - c0 * n12;
        c1 = (int)(((float)r + 0.5f) * inv2);
        c2 = r - c1 * n2;
    }
    __device__ __forceinline__ int index(int c0, int c1, int c2) const { return c2 + n2 * (c1 + n1 * c0); }
    // x in (-n, 2n) -> x mod n
    __device__ __forceinline__ static int wrap(int x, int n)
    {
        x = x < 0 ? x + n : x;
        return x >= n ? x - n : x;
    }
    __device__ __forceinline__ int sum(int a0, int a1, int a2, int b0, int b1, int b2) const  // q1 + q2
    {
        return index(wrap(a0 + b0 - h0, n0), wrap(a1 + b1 - h1, n1), wrap(a2 + b2 - h2, n2));
    }
    __device__ __forceinline__ int diff(int a0, int a1, int a2, int b0, int b1, int b2) const  // q1 - q2
    {
        return index(wrap(a0 - b0 + h0, n0), wrap(a1 - b1 + h1, n1), wrap(a2 - b2 + h2, n2));
    }
    __device__ __forceinline__ int neg_sum(int a0, int a1, int a2, int b0, int b1, int b2) const  // -q1 - q2
    {
        return index(wrap(3 * h0 - a0 - b0, n0), wrap(3 * h1 - a1 - b1, n1), wrap(3 * h2 - a2 - b2, n2));
    }
    __device__ __forceinline__ int neg(int b0, int b1, int b2) const  // -q
    {
        return index(b0 == 0 ? 0 : n0 - b0, b1 == 0 ? 0 : n1 - b1, b2 == 0 ? 0 : n2 - b2);
    }
    __device__ __forceinline__ bool sum_inside(int a0, int a1, int a2, int b0, int b1, int b2) const
    {
        int s0 = a0 + b0, s1 = a1 + b1, s2 = a2 + b2;
        return s0 >= h0 && s0 < 3 * h0 && s1 >= h1 && s1 < 3 * h1 && s2 >= h2 && s2 < 3 * h2;
    }
    __device__ __forceinline__ bool diff_inside(int a0, int a1, int a2, int b0, int b1, int b2) const
    {
        int t0 = a0 - b0, t1 = a1 - b1, t2 = a2 - b2;
        return t0 >= -h0 && t0 < h0 && t1 >= -h1 && t1 < h1 && t2 >= -h2 && t2 < h2;
    }
    __device__ __forceinline__ int gamma() const { return index(h0, h1, h2); }
};

}  // namespace alt3
