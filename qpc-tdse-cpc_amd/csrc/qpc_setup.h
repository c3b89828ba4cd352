// qpc_setup.h -- internal header of the host-side set-up (public API in include/qpc_gpu.h).
#pragma once
#include "../../include/qpc_gpu.h"
#include "qpc_common.h"
