// Internal host/device structures shared by the kernels and the C ABI: the
// memory layout (tf_layout.h) and the device arithmetic (tf_device.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tf_layout.h"
#include "tf_device.h"
