// Target of the device code: HIP on CDNA4 (gfx950), or, when TGR_WAVE_EMU
// is defined, the CPU wave emulator of the test suite (tests/emu), which runs
// the same solver source with one host thread per lane.
#pragma once
#ifdef TGR_WAVE_EMU
#include "wave_emu.h"
#define TGR_OPAQUE_PTR "+r"
#else
#include <hip/hip_runtime.h>
#define TGR_OPAQUE_PTR "+s"
#endif
