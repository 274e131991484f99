// Shared declarations of the C-API implementation files.
#pragma once

#include "fft_launch.h"
#include "grf.h"
#include "runtime.h"

namespace mitos {

int set_last_error(int e);


// k-space kernel selected from a reference callback address
struct KSpec {
    int id;        // KID_*
    double p[4];   // parameters (see ksp_value)
};
enum {
    KID_CONSTANT = 0,
    KID_INV_POISSON = 1,
    KID_DX = 2,
    KID_DY = 3,
    KID_DZ = 4,
    KID_POWER_NO_TRANSFER = 5,
    KID_UNDO_WINDOW = 6,
    KID_SQRT_POWER_LAW = 7,
    KID_BBKS = 8,
    KID_LOWPASS = 9,
    KID_HIPASS = 10,
    KID_TOPHAT = 11,
    KID_ELLIPTIC_TOPHAT = 12,
    KID_TRANSLATE = 13,
    KID_GAUSSIAN = 14,
    KID_GAUSSIAN_INV = 15,
};


}  // namespace mitos
