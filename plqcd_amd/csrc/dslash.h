// Launch parameters of the hopping kernels.
#pragma once

#include "device_common.h"

#ifndef HOP_BLOCK
#define HOP_BLOCK 256
#endif

struct HopArgs {
   int L[4];     // extents x, y, z, t
   int Lxh;      // L[0] / 2: sites of one parity in an x row
   int Vh;       // sites per parity
   int nparity;  // 1 (eo or oe) or 2 (full H)
   int parity0;  // output parity when nparity == 1
   size_t rhs_stride;  // elements of the storage type between RHS
};

HopArgs make_hop_args(const plqcd_hip_lattice *lat, int parity);
