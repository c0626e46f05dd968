// Stand-in <hip/hip_runtime.h> for the CPU build of the library's HIP sources
// (tests only): everything comes from the emulator header.
#pragma once
#include "../../hipemu.hpp"
