/* QLL API, double precision (D suffix): device fields and links in FP64. */
#define QLL_PREC_DOUBLE
#include "qll_prec.inc"
