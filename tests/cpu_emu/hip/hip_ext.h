// CPU emulation: hipExtLaunchKernelGGL is declared with the rest of the runtime
#pragma once
#include "hip_runtime.h"
