// x passes (1 and 5) for double plans.
#define PPF_T double
#include "ppf_xpass.inc"
