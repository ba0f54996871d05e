/*
 * gsis_oracle_ns.c -- CPU oracle of the synthetic NS solve (mode M1).
 * TEST INFRASTRUCTURE ONLY (see gsis_oracle.c).
 */
#include "gsis_oracle.h"
int orc_synthetic_ns(orc_t *o) { (void)o; return GSIS_EUNSUPPORTED; }
