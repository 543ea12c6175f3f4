/* Exhaustive check of the division-free x / 4294967295 used by the McICA
 * kernel's conversion of MT19937 draws: for every 32-bit draw y, both the
 * direct value y and the default-REAL-rounded value of non-negative draws,
 * the FMA form must equal the IEEE quotient.  Prints the mismatch count. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>

static double div_2p32m1(double x)
{
    const double S = 2.3283064365386963e-10;   /* 2^-32 */
    const double v = x * S;
    const double q0 = fma(v, S, v);
    const double rem = fma(-q0, 4294967295.0, x);
    return fma(rem, 1.0 / 4294967295.0, q0);
}

int main(void)
{
    unsigned long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long yy = 0; yy < 4294967296LL; ++yy) {
        const uint32_t y = (uint32_t)yy;
        const int32_t ys = (int32_t)y;
        const double x = ys < 0 ? (double)ys + 4294967296.0 : (double)(float)ys;
        if (x / 4294967295.0 != div_2p32m1(x)) ++bad;
        if ((double)y / 4294967295.0 != div_2p32m1((double)y)) ++bad;
    }
    printf("%llu\n", bad);
    return 0;
}
