// Resident device data handles shared by the C ABI (capi.hip) and the
// request server (server.hip).
#pragma once
#include <stdint.h>

struct gm_bases {
    int curve;
    int n;
    int dev;
    void* d;  // Aff<O>[n] in the engine's resident format
};

struct gm_scalars {
    int n;
    int limbs;
    int dev;
    uint64_t* d;
};
