// CPU emulation: the runtime API is declared with the rest of the runtime
#pragma once
#include "hip_runtime.h"
