// x passes (1 and 5) of float plans, line lengths of group A.
#define PPF_T float
#define PPF_GROUP PPF_SIZES_A
#include "ppf_xpass.inc"
