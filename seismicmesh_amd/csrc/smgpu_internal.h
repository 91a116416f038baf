// Internal declarations shared by the library's translation units.
#pragma once
#include <cstdint>

#include "../../include/smgpu.h"

// Records a message for sm_last_error() and returns `code`.
int sm_set_error(int code, const char *fmt, ...);
