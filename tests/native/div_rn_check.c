/* The FP64 kernels divide by launch constants as q = x*r, e = fma(-q, c, x),
 * q' = fma(e, r, q) with r = RN(1/c) (hedgehog_lpbf_amd/csrc/lpbf_f64.hip,
 * div_rn).  This checks q' == x / c (IEEE, correctly rounded) for the
 * kernels' divisors (rho_cp, rho_cp*h, dSub, the IN718 table steps, the
 * mushy-zone width, rho*cp_eff over its range) and for random significands,
 * over operands spanning many binades.  Prints the mismatch count.  With the
 * argument "naive" it checks q = x*r alone, which must mismatch. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static int naive = 0;

static double div_rn(double x, double c, double r) {
    const double q = x * r;
    if (naive) return q;
    const double e = fma(-q, c, x);
    return fma(e, r, q);
}

static uint64_t s = 88172645463325252ull;
static uint64_t next(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double unit(void) { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

int main(int argc, char** argv) {
    naive = argc > 1;
    (void)argv;
    const double rho = 8190.0, cp = 565.0, h = 40e-6, dsub = 30e-6;
    const double T[9] = {298, 473, 673, 873, 1073, 1273, 1473, 1533, 1609};
    double cs[16];
    int nc = 0;
    cs[nc++] = rho * cp;
    cs[nc++] = rho * cp * h;
    cs[nc++] = dsub;
    cs[nc++] = 1609.0 - 1533.0;
    for (int i = 1; i < 9; ++i) cs[nc++] = T[i] - T[i - 1];
    long bad = 0, tot = 0;
    for (int ci = 0; ci < nc; ++ci) {
        const double c = cs[ci], r = 1.0 / c;
        for (long n = 0; n < 2000000; ++n) {
            const double x = (unit() * 2 - 1) * ldexp(1.0, (int)(next() % 200) - 100);
            bad += (x / c != div_rn(x, c, r));
            ++tot;
        }
    }
    for (long m = 0; m < 20000; ++m) {   /* rho * cp_eff(T), cp_eff in [400, 3900] */
        const double c = rho * (400.0 + unit() * 3500.0), r = 1.0 / c;
        for (int n = 0; n < 200; ++n) {
            const double x = (unit() * 2 - 1) * ldexp(1.0, (int)(next() % 120) - 60);
            bad += (x / c != div_rn(x, c, r));
            ++tot;
        }
    }
    for (long m = 0; m < 200000; ++m) {   /* random significands */
        const double c = (1 + unit()) * ldexp(1.0, (int)(next() % 60) - 30), r = 1.0 / c;
        for (int n = 0; n < 20; ++n) {
            const double x = (1 + unit()) * ldexp(1.0, (int)(next() % 60) - 30);
            bad += (x / c != div_rn(x, c, r));
            ++tot;
        }
    }
    printf("%ld %ld\n", bad, tot);
    return bad != 0;
}
