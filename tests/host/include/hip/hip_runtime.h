#pragma once
// Host build of the walker's device functions (tests only): the few HIP
// names siddon_ray.h uses, as plain C++.
#include <math.h>
#include <stdint.h>
#include <string.h>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
struct float4 {
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int __float_as_int(float f) {
    int i;
    memcpy(&i, &f, 4);
    return i;
}
inline float __int_as_float(int i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
}
unsigned long long __ballot(int);  // declared for walk_dispatch; not called here
