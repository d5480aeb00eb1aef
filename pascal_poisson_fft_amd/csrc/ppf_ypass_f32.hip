// y passes (2 and 4) for float plans.
#define PPF_T float
#include "ppf_ypass.inc"
