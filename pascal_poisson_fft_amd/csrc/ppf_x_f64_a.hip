// x passes (1 and 5) of double plans, line lengths of group A.
#define PPF_T double
#define PPF_GROUP PPF_SIZES_A
#include "ppf_xpass.inc"
