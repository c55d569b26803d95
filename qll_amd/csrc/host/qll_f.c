/* QLL API, single precision (F suffix): device fields and links in FP32. */
#define QLL_PREC_FLOAT
#include "qll_prec.inc"
