// dmath.h -- float vector, color and quaternion math for the HIP kernels and
// their host-side setup code.
//
// The operators keep fresnel's float evaluation order, because primary rays
// and hit ids must match the CPU path bit for bit (the kernels are compiled
// with -ffp-contract=off and IEEE division and square root):
//   * dot products sum left to right            (fresnel/common/VectorMath.h:570-573)
//   * `vec / scalar` multiplies by 1/scalar     (VectorMath.h:508-512, ColorMath.h:215-219)
//   * `vec /= scalar` divides each component    (VectorMath.h:536-542, ColorMath.h:243-249)
//   * fast::rsqrt(x) is 1.0f / sqrtf(x)         (VectorMath.h:46-53, host branch)
//   * quaternion rotation uses the bkph form    (VectorMath.h:1080-1092)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define FH_HD __host__ __device__ __forceinline__

namespace fh
{
struct v3
    {
    float x, y, z;
    };
struct v2
    {
    float x, y;
    };
struct quat
    {
    float s;
    v3 v;
    };

FH_HD v3 mk3(float x, float y, float z)
    {
    v3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
    }
FH_HD v3 operator+(v3 a, v3 b)
    {
    return mk3(a.x + b.x, a.y + b.y, a.z + b.z);
    }
FH_HD v3 operator-(v3 a, v3 b)
    {
    return mk3(a.x - b.x, a.y - b.y, a.z - b.z);
    }
FH_HD v3 operator-(v3 a)
    {
    return mk3(-a.x, -a.y, -a.z);
    }
FH_HD v3 operator*(v3 a, v3 b)
    {
    return mk3(a.x * b.x, a.y * b.y, a.z * b.z);
    }
FH_HD v3 operator*(v3 a, float s)
    {
    return mk3(a.x * s, a.y * s, a.z * s);
    }
FH_HD v3 operator*(float s, v3 a)
    {
    return mk3(a.x * s, a.y * s, a.z * s);
    }
// `vec / scalar`: multiply by the rounded reciprocal
FH_HD v3 div_rcp(v3 a, float s)
    {
    const float q = 1.0f / s;
    return a * q;
    }
// `vec /= scalar`: true division per component
FH_HD v3 div_true(v3 a, float s)
    {
    return mk3(a.x / s, a.y / s, a.z / s);
    }
FH_HD float dot(v3 a, v3 b)
    {
    return a.x * b.x + a.y * b.y + a.z * b.z;
    }
FH_HD v3 cross(v3 a, v3 b)
    {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
    }
FH_HD v2 operator-(v2 a, v2 b)
    {
    v2 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
    }
FH_HD v2 operator+(v2 a, v2 b)
    {
    v2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
    }
FH_HD v2 operator*(float s, v2 a)
    {
    v2 r;
    r.x = a.x * s;
    r.y = a.y * s;
    return r;
    }
FH_HD float dot(v2 a, v2 b)
    {
    return a.x * b.x + a.y * b.y;
    }
FH_HD quat conj(quat a)
    {
    quat r;
    r.s = a.s;
    r.v = -a.v;
    return r;
    }
FH_HD v3 rotate(quat a, v3 b)
    {
    return (a.s * a.s - dot(a.v, a.v)) * b + 2 * a.s * cross(a.v, b) + 2 * dot(a.v, b) * a.v;
    }
// (1 - x) a + x b, fresnel/common/ColorMath.h:504-507
FH_HD v3 lerp(float x, v3 a, v3 b)
    {
    return (1.0f - x) * a + x * b;
    }
// std::min / std::max semantics: (b < a) ? b : a (fresnel's fast::min/max on the host)
FH_HD float smin(float a, float b)
    {
    return (b < a) ? b : a;
    }
FH_HD float smax(float a, float b)
    {
    return (a < b) ? b : a;
    }

// ------------------------------------------------------------ division by a constant
// n / d for a divisor d >= 1 fixed per image, by a multiply-high and two
// shifts: the round-up method of Granlund and Montgomery ("Division by
// invariant integers using multiplication", PLDI 1994, section 4) with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1, exact for every
// 32-bit n.
struct FastDiv
    {
    uint32_t m, s1, s2;
    };

inline FastDiv make_fastdiv(uint32_t d)
    {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d)
        l++;
    FastDiv f;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.s1 = l ? 1u : 0u;
    f.s2 = l ? l - 1 : 0u;
    return f;
    }

FH_HD uint32_t fdiv(uint32_t n, FastDiv f)
    {
    const uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
    return (t + ((n - t) >> f.s1)) >> f.s2;
    }

// ------------------------------------------------------------ Philox4x32-10
// Counter-based RNG of Salmon, Moraes, Dror and Shaw (SC'11), 10 rounds,
// multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl increments 0x9E3779B9 /
// 0xBB67AE85, as fresnel calls it through Random123 (fresnel/common/RayGen.h:77-79).
struct u4
    {
    uint32_t v[4];
    };

// a ^ b ^ c; one v_bitop3_b32 (truth table 0x96) on gfx950
FH_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
    {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
    }

FH_HD u4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
    {
#pragma unroll
    for (int r = 0; r < 10; r++)
        {
        // one 32x32->64 multiply gives both halves (v_mad_u64_u32 on the device)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32);
        const uint32_t lo1 = (uint32_t)p1;
        const uint32_t n0 = xor3(hi1, c1, k0);
        const uint32_t n2 = xor3(hi0, c3, k1);
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
        }
    u4 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
    }

// Random123 u01<float>: float(u) * 2^-32 + 2^-33, in (0, 1]
FH_HD float u01(uint32_t u)
    {
    return (float)u * 0x1p-32f + 0x1p-33f;
    }

// purpose constants of the fourth counter word
constexpr uint32_t RNG_MIS = 0x8754abcdu;      // fresnel/common/RayGen.h:34
constexpr uint32_t RNG_RR = 0x54abf853u;       // RayGen.h:37
constexpr uint32_t RNG_AA = 0x22ab5871u;       // fresnel/common/Camera.h:305
constexpr uint32_t RNG_APERTURE = 0x983abc12u; // Camera.h:302

constexpr float PI_F = 3.14159265358979323846f; // float(M_PI)
    } // namespace fh
