// y passes (2 and 4) of double plans, line lengths of group A.
#define PPF_T double
#define PPF_GROUP PPF_SIZES_A
#include "ppf_ypass.inc"
