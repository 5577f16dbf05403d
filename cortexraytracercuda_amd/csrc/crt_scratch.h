// crt_scratch.h - size and layout limits of the no-allocation render's device scratch array
// (noalloc.hip), shared by the array's definition and the host code that carves it up.
#pragma once
#include <stddef.h>
#include <stdint.h>

// 128 MiB per GPU: 256K live path slots (164 B each), 1M path ids per pass (16 B of radiance each),
// closest-hit overflow stacks for 2 workgroups per CU on 256 CUs (384 B per thread), control block,
// counters and a two-entry request table.
#define kCrtScratchBytes (128ull << 20)
#define kCrtScratchSlots (1u << 18)
#define kCrtScratchPassIds (1u << 20)
#define kCrtScratchExtendPerCu 2u

#ifdef __cplusplus
extern "C" {
#endif
// Device address (and size) of the scratch array on `device`, which must be current. Resolved with
// hipGetSymbolAddress on the first call per device and kept. Returns a hipError_t.
int crt_scratch_setup(int device, void** ptr, size_t* bytes);
#ifdef __cplusplus
}
#endif
