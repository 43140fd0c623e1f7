/* Accuracy of mdt_exp (md_tian2_amd/csrc/mdt_math.h) against libm exp:
 * prints the maximum error in ulp over a dense sweep of [-700, 700] and a
 * fine sweep of the EMT argument range [-120, 70]. */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include "../../md_tian2_amd/csrc/mdt_math.h"

static double ulp_err(double a, double b) {
    double u = nextafter(fabs(b), INFINITY) - fabs(b);
    return fabs(a - b) / u;
}

int main(void) {
    static const double tab[MDT_EXP_TAB_N] = MDT_EXP_TAB_INIT;
    double worst = 0.0, wx = 0.0;
    for (long i = 0; i <= 4000000; ++i) {
        const double x = -700.0 + 1400.0 * (double)i / 4000000.0;
        const double e = ulp_err(mdt_exp(x, tab), exp(x));
        if (e > worst) { worst = e; wx = x; }
    }
    for (long i = 0; i <= 4000000; ++i) {
        const double x = -120.0 + 190.0 * (double)i / 4000000.0 + 1e-7 * (double)(i % 7);
        const double e = ulp_err(mdt_exp(x, tab), exp(x));
        if (e > worst) { worst = e; wx = x; }
    }
    printf("%.6f %.17g\n", worst, wx);
    return 0;
}
