// y passes (2 and 4) for double plans.
#define PPF_T double
#include "ppf_ypass.inc"
