/* The oracle compiled with operation counting (see opcount.hpp): test and
 * measurement infrastructure.  Exports the oracle's own entry points (ora_*)
 * plus ora_opcounts_reset / ora_opcounts_get. */
#include "opcount.hpp"

extern "C" {
OpCounts g_opcounts;

void ora_opcounts_reset(void) { memset(&g_opcounts, 0, sizeof(g_opcounts)); }

/* add (incl. subtract), mul, div, sqrt, neg, fabs, cmp */
void ora_opcounts_get(unsigned long long out[7])
{
  out[0] = g_opcounts.add;
  out[1] = g_opcounts.mul;
  out[2] = g_opcounts.div;
  out[3] = g_opcounts.sqrt;
  out[4] = g_opcounts.neg;
  out[5] = g_opcounts.fabs;
  out[6] = g_opcounts.cmp;
}

#include "mhd_oracle.c"
}
