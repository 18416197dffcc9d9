// sincos_poly.h -- double-precision sine and cosine of a float in [0, 8] with a
// rounding test, shared by the device shading code (rt/shade.h) and the host
// program that tabulates the arguments whose rounding the test cannot decide
// (rt/gen_sincos_hard.cpp), so both evaluate bit-identical doubles.
//
// x is reduced by k pi/2 (pi/2 split in two doubles, fused multiply-adds: the
// reduced argument is accurate to about 2^-52 relative, as no float in the
// range other than 0 lies closer than 4e-8 to a multiple of pi/2), and sin and
// cos of the reduced argument (|r| <= pi/4) are Taylor polynomials through
// r^15 and r^16 (truncation below 2^-53).  The result rounds correctly to
// float when every value within a relative 2^-46 of the double result rounds
// to the same float; otherwise the argument is "hard".
#pragma once

#ifndef FH_SC_QUAL
#define FH_SC_QUAL static inline
#endif

namespace fh
{
// Double sine and cosine of x; returns true when either cannot be rounded to
// float from them with certainty.
FH_SC_QUAL bool sincos_poly(float x, double& sd, double& cd)
    {
    const double xd = x;
    const double k = __builtin_rint(xd * 0.63661977236758134308);
    double r = __builtin_fma(-k, 1.5707963267948966192, xd);
    r = __builtin_fma(-k, 6.1232339957367660359e-17, r);
    const int q = (int)k & 3;
    const double r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - ... - r^14/15!)
    double ps = -7.6471637318198164759e-13;
    ps = __builtin_fma(ps, r2, 1.6059043836821614599e-10);
    ps = __builtin_fma(ps, r2, -2.5052108385441718775e-08);
    ps = __builtin_fma(ps, r2, 2.7557319223985890653e-06);
    ps = __builtin_fma(ps, r2, -1.9841269841269841270e-04);
    ps = __builtin_fma(ps, r2, 8.3333333333333333333e-03);
    ps = __builtin_fma(ps, r2, -1.6666666666666666667e-01);
    const double sr = __builtin_fma(ps * r2, r, r);
    // cos r = 1 - r^2/2! + r^4/4! - ... + r^16/16!
    double pc = 4.7794773323873852974e-14;
    pc = __builtin_fma(pc, r2, -1.1470745597729724714e-11);
    pc = __builtin_fma(pc, r2, 2.0876756987868098979e-09);
    pc = __builtin_fma(pc, r2, -2.7557319223985890653e-07);
    pc = __builtin_fma(pc, r2, 2.4801587301587301587e-05);
    pc = __builtin_fma(pc, r2, -1.3888888888888888889e-03);
    pc = __builtin_fma(pc, r2, 4.1666666666666666667e-02);
    pc = __builtin_fma(pc, r2, -0.5);
    const double cr = __builtin_fma(pc, r2, 1.0);
    sd = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
    cd = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
    const double lo = 1.0 - 0x1p-46, hi = 1.0 + 0x1p-46;
    return (float)(sd * lo) != (float)(sd * hi) || (float)(cd * lo) != (float)(cd * hi);
    }
    } // namespace fh
