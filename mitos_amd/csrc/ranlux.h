// RANLUX in double precision (Luescher, Comput. Phys. Commun. 79 (1994) 100;
// the 48-bit variant "ranlxd"), luxury level 1 (202 steps per 12 delivered
// numbers), as the GSL generator gsl_rng_ranlxd1 that src/grf_ngeniclike.c
// uses: the subtract-with-borrow recurrence
//     x_t = x_{t-5} - x_{t-12} - c_{t-1}  (mod 1),  c_t = 1 (in units of 2^-48)
//     when the difference before the wrap was negative,
// on twelve 48-bit fractions held exactly in doubles.  Each refill advances
// the recurrence p = 202 steps and the next twelve outputs are the last
// twelve values, oldest first.  A seed s (0 -> 1) initialises the twelve
// values from the bits of a 31-bit shift register (taps 0 and 18) loaded with
// the low 31 bits of s.  Every operation is exact in double precision, so the
// host and device streams are identical.
#pragma once

#include <hip/hip_runtime.h>

namespace mitos {

struct Ranlxd {
    double x[12];
    double carry;
    int ir, jr, ir_old, pr;
};

constexpr double kRanlxdBit = 1.0 / 281474976710656.0;  // 2^-48

__host__ __device__ inline void ranlxd_seed(Ranlxd& s, unsigned long seed, int luxury = 202) {
    if (seed == 0) seed = 1;
    int v = (int)(seed & 0xFFFFFFFFUL);  // low bits of the seed
    int bits[31];
    for (int k = 0; k < 31; k++) {
        bits[k] = v % 2;
        v /= 2;
    }
    int ib = 0, jb = 18;
    for (int k = 0; k < 12; k++) {
        double a = 0;
        for (int l = 0; l < 48; l++) {
            a += a + (double)((bits[ib] + 1) % 2);
            bits[ib] = (bits[ib] + bits[jb]) % 2;
            ib = ib == 30 ? 0 : ib + 1;
            jb = jb == 30 ? 0 : jb + 1;
        }
        s.x[k] = kRanlxdBit * a;
    }
    s.carry = 0;
    s.ir = 11;
    s.jr = 7;
    s.ir_old = 0;
    s.pr = luxury;
}

// p steps of the recurrence from slot ir (slot jr = ir - 5 holds x_{t-5})
__host__ __device__ inline void ranlxd_refill(Ranlxd& s) {
#pragma clang fp contract(off)
    int ir = s.ir, jr = s.jr;
    double c = s.carry;
    for (int k = 0; k < s.pr; k++) {
        double y = s.x[jr] - s.x[ir] - c;
        if (y < 0) {
            c = kRanlxdBit;
            y += 1;
        } else {
            c = 0;
        }
        s.x[ir] = y;
        ir = ir == 11 ? 0 : ir + 1;
        jr = jr == 11 ? 0 : jr + 1;
    }
    s.ir = ir;
    s.ir_old = ir;
    s.jr = jr;
    s.carry = c;
}

// the next number in [0, 1) (48-bit resolution)
__host__ __device__ inline double ranlxd_uniform(Ranlxd& s) {
    s.ir = s.ir == 11 ? 0 : s.ir + 1;
    if (s.ir == s.ir_old) ranlxd_refill(s);
    return s.x[s.ir];
}

}  // namespace mitos
