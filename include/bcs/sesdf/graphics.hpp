// Outputs of the complete-surface builder, with the field names of the reference's
// bcs::sesdf::SesdfGraphics / SesdfData (sesdf/include/bcs/sesdf/graphics.hpp:10-66).
// Device pointers use HIP's vector types; compile with hipcc, or with a host compiler and
// -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include.
#ifndef BCS_SESDF_GRAPHICS_HPP
#define BCS_SESDF_GRAPHICS_HPP

#include <hip/hip_vector_types.h>

#include "bcs/core/type.hpp"

using GLuint = unsigned int;

namespace bcs::sesdf
{
    // No OpenGL interop: ids are 0, offsets and sizes (bytes) describe the device arrays.
    struct SesdfGraphics
    {
        uint32_t           atomNb = 0;
        HandleSpan<GLuint> atoms;

        uint32_t           convexPatchNb = 0;
        HandleSpan<GLuint> convexPatches;

        uint32_t           sectorNb = 0;
        HandleSpan<GLuint> sectors;

        uint32_t           segmentPatchNb = 0;
        HandleSpan<GLuint> segmentPatches;

        uint32_t           circlePatchNb = 0;
        HandleSpan<GLuint> circlePatches;

        uint32_t           concavePatchNb = 0;
        HandleSpan<GLuint> concavePatchesPosition;
        HandleSpan<GLuint> concavePatchesId;
        HandleSpan<GLuint> concavePatchesNeighbors;

        float    probeRadius         = 0.f;
        uint16_t maxConcaveNeighbors = 0;
    };

    struct SesdfData
    {
        uint32_t atomNb = 0;
        float4 * atoms  = nullptr; // grid-sorted (xyz, vdW radius)

        uint32_t convexPatchNb = 0;
        uint2 *  convexPatches = nullptr; // per atom [start, end) into sectors

        uint32_t sectorNb = 0;
        float4 * sectors  = nullptr; // (axis, half angle)

        uint32_t segmentPatchNb = 0;
        uint4 *  segmentPatches = nullptr; // (i, j, intersection A, intersection B)

        uint32_t circlePatchNb = 0;
        uint2 *  circlePatches = nullptr; // full circles (i, j)

        uint32_t concavePatchNb          = 0;
        float4 * concavePatchesPosition  = nullptr; // (probe xyz, +-1 / +-2)
        int4 *   concavePatchesId        = nullptr; // (i, j, k, neighbour count)
        float4 * concavePatchesNeighbors = nullptr; // maxConcaveNeighbors per with-neighbour probe

        float    probeRadius         = 0.f;
        uint16_t maxConcaveNeighbors = 0;
    };
} // namespace bcs::sesdf

#endif // BCS_SESDF_GRAPHICS_HPP
