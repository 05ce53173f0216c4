/* libm_interpose.c -- test harness only.  Linked into oracle/_ref/libhvcref_ora.so
 * (with -Wl,-Bsymbolic-functions) so that the reference classes' calls to exp
 * and log resolve to the oracle's functions instead of libm's. */
#include "../hvc_oracle.h"
double exp(double x) { return ora_exp(x); }
double log(double x) { return ora_log(x); }
