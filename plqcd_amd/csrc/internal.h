// Internal declarations shared by the library's translation units.
#pragma once

#include <cstddef>
#include <cstdint>
#include "../../include/plqcd_hip.h"

namespace plqcd {

// Records a formatted error message for plqcd_hip_last_error(); returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Host <-> device transfer of the sites [first, last) (even-odd indices) of
// one RHS; synchronous with respect to the host buffer.
int field_upload_range(plqcd_hip_field *f, int rhs, const spinor *host, int first, int last);
int field_download_range(const plqcd_hip_field *f, int rhs, spinor *host, int first, int last);
// Device copy of one parity of one RHS (F64/F32 fields), asynchronous.
int field_copy_range(const plqcd_hip_field *src, plqcd_hip_field *dst, int rhs, int parity);

// Releases the device state owned by the drop-in entry points (dirac_api).
void release_dropin_state();

// memcpy of host memory, split over a small pool of threads when large.
void host_copy(void *dst, const void *src, size_t n);

}  // namespace plqcd
