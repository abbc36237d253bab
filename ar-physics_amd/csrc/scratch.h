// Fixed per-GPU device scratch for the entry points that allocate nothing (scratch.hip).
#pragma once
#include <stddef.h>

// 128 MiB per GPU: the two chunk partials of a 6-10 term lincomb of 1,048,576 particles
// (2 x 48 MiB) with room for solids.
#define kArpScratchBytes ((size_t)128 << 20)
#define kArpScratchMaxDevices 64

#ifdef __cplusplus
extern "C" {
#endif
int arp_scratch_acquire(int device, void** ptr, size_t* bytes);
#ifdef __cplusplus
}
#endif
