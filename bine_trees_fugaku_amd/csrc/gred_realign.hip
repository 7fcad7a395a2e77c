// gred_realign.hip -- the realigning instantiations of gred_kernel (REALIGN =
// true: outputs that start off a 16-byte boundary are stored as aligned chunks),
// compiled as a unit of their own so the two halves build in parallel.
#define BINE_GRED_REALIGN 1
#include "gred.hip"
