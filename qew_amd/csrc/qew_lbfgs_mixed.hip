// The batched L-BFGS kernels for preconditioned lines whose half-length has
// factors 3, 5 or 7 (the radix-3/5/7 LDS FFT stages), compiled as a unit of
// their own from the same source: instantiated beside the power-of-two
// kernels, they change how the compiler inlines the device code all of them
// share, and the power-of-two kernels then spill more registers.
#define QEW_LBFGS_MIXED_UNIT
#include "qew_lbfgs_batch.hip"
