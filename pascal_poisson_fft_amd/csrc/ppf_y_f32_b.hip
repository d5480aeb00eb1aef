// y passes (2 and 4) of float plans, line lengths of group B.
#define PPF_T float
#define PPF_GROUP PPF_SIZES_B
#include "ppf_ypass.inc"
