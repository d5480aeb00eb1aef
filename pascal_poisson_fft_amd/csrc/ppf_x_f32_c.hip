// x passes (1 and 5) of float plans, line lengths of group C.
#define PPF_T float
#define PPF_GROUP PPF_SIZES_C
#include "ppf_xpass.inc"
