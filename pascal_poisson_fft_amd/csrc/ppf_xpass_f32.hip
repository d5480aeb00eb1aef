// x passes (1 and 5) for float plans.
#define PPF_T float
#include "ppf_xpass.inc"
