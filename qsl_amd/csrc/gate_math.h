// gate_math.h -- the per-amplitude gate arithmetic (gate_core.h) with the
// library's host-side declarations.
#pragma once

#include "internal.h"
#include "gate_core.h"
