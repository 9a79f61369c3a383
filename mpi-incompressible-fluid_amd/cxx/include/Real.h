// Floating-point type of the host-side interface (include/Real.h of the
// reference, built there with USE_DOUBLE).  Host arrays exchanged with the
// device library are always double; the device context decides whether the
// fields are stored in FP64 or in the mixed FP32 form.
#ifndef MIF_SHIM_REAL_H
#define MIF_SHIM_REAL_H

#ifndef Real
#define Real double
#endif

#endif  // MIF_SHIM_REAL_H
