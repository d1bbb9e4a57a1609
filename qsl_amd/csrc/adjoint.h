// adjoint.h -- generators of the parametrised gates, shared by the
// per-circuit adjoint path (adjoint.hip) and the batched one (batch.hip).
//
// A parametrised gate is G(theta) = exp(-i theta P / 2); its gradient term
// is Im <lambda| P |psi> (rotations: P = X, Y or Z on the target, times the
// control projector for the controlled forms; phase / controlPhase:
// P = -2 |1><1| on the target; the number-preserving rotations act on the
// {01, 10} pair of their two qubits).
#pragma once

#include "qslhip.h"

namespace qslhip {

enum GenKind { GEN_NONE = 0, GEN_X, GEN_Y, GEN_Z, GEN_P1, GEN_NPX, GEN_NPY };

__host__ __device__ inline int gen_of(int kind)
{
    switch (kind) {
    case QSLHIP_ROT_X: case QSLHIP_CROT_X: return GEN_X;
    case QSLHIP_ROT_Y: case QSLHIP_CROT_Y: return GEN_Y;
    case QSLHIP_ROT_Z: case QSLHIP_CROT_Z: return GEN_Z;
    case QSLHIP_PHASE: case QSLHIP_CPHASE: return GEN_P1;
    case QSLHIP_NP_ROT_X: return GEN_NPX;
    case QSLHIP_NP_ROT_Y: return GEN_NPY;
    default: return GEN_NONE;
    }
}

}  // namespace qslhip
