// F16 spinors with 18-real links: instantiation of the hopping kernels.
#include "dslash_kernels.h"

template int plqcd::hop_launch<PLQCD_PREC_F16, 18>(const plqcd_hip_gauge *, const plqcd_hip_field *, plqcd_hip_field *,
                                                  int, int, const plqcd::DslashOp &);
template int plqcd::hop_batch_launch<PLQCD_PREC_F16, 18>(int, const plqcd_hip_gauge *const *,
                                                        const plqcd_hip_field *const *, plqcd_hip_field *const *, int);
