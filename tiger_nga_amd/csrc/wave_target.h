// Target of the device code: HIP on CDNA4 (gfx950), or, when TGR_WAVE_EMU
// is defined, the CPU wave emulator of the test suite (tests/emu), which runs
// the same solver source with one host thread per lane.
#pragma once
#ifdef TGR_WAVE_EMU
#include "wave_emu.h"
#define TGR_OPAQUE_PTR "+r"
#define TGR_GLOBAL
#define TGR_LDS
#define TGR_CONST
#else
#include <hip/hip_runtime.h>
#define TGR_OPAQUE_PTR "+s"
// pointee qualifier of device (HBM) memory: loads through it are global_load,
// not flat_load
#define TGR_GLOBAL __attribute__((address_space(1)))
// pointee qualifier of LDS: ds_read / ds_write instead of flat accesses
#define TGR_LDS __attribute__((address_space(3)))
// pointee qualifier of read-only device memory (the constant address space):
// loads through it with a wave-uniform address are scalar loads
#define TGR_CONST __attribute__((address_space(4)))
#endif
