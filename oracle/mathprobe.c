/* TEST INFRASTRUCTURE ONLY: exposes the portable elementary functions of
 * ocotillo_amd/csrc/oco_math.h (host build) to the tests through ctypes. */
#include <string.h>
#include "../ocotillo_amd/csrc/oco_math.h"

void mp_eval(int fn, long n, const double* x, const double* y, double* out)
{
    for (long i = 0; i < n; i++) {
        switch (fn) {
        case 0: out[i] = oco_exp(x[i]); break;
        case 1: out[i] = oco_log(x[i]); break;
        case 2: out[i] = oco_log10(x[i]); break;
        case 3: out[i] = oco_exp10(x[i]); break;
        case 4: out[i] = oco_pow(x[i], y[i]); break;
        default: out[i] = oco_sqrt(x[i]); break;
        }
    }
}
