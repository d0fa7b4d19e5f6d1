/* Prints sizeof(sg_config) and offsets of the members that
 * fortran_layout_check.f90 prints, for tests/test_fortran_binding.py. */
#include <stddef.h>
#include <stdio.h>
#include "senga_gpu.h"
#define P(m) printf(#m " %zu\n", offsetof(sg_config, m))
int main(void) {
  printf("size %zu\n", sizeof(sg_config));
  P(xgdlen); P(bc_rpar); P(yrin); P(ncpoly); P(amascp); P(rparam); P(mslist);
  P(effy3b); P(clnten); P(ncovis); P(diffco); P(wilko2);
  return 0;
}
