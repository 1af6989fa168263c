// Outputs of the exterior-surface builder, with the field names of the reference's
// bcs::ssesdf::SsesdfGraphics / SsesdfData (sesdf/include/bcs/ssesdf/graphics.hpp:10-50).
// Device pointers use HIP's vector types; compile with hipcc, or with a host compiler and
// -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include.
#ifndef BCS_SSESDF_GRAPHICS_HPP
#define BCS_SSESDF_GRAPHICS_HPP

#include <hip/hip_vector_types.h>

#include "bcs/core/type.hpp"

using GLuint = unsigned int;

namespace bcs::ssesdf
{
    // No OpenGL interop: ids are 0, offsets and sizes (bytes) describe the device arrays.
    struct SsesdfGraphics
    {
        uint32_t           atomNb = 0;
        HandleSpan<GLuint> atoms;

        uint32_t           circlePatchNb = 0;
        HandleSpan<GLuint> circlePosition;
        HandleSpan<GLuint> circleAxis;
        HandleSpan<GLuint> circleVs;

        uint32_t           concavePatchNb = 0;
        HandleSpan<GLuint> concavePatchesPosition;
        HandleSpan<GLuint> concavePatchesId;
        HandleSpan<GLuint> concavePatchesNeighbors;

        float    probeRadius         = 0.f;
        uint16_t maxConcaveNeighbors = 0;
    };

    struct SsesdfData
    {
        uint32_t atomNb = 0;
        float4 * atoms  = nullptr; // grid-sorted (xyz, vdW radius)

        uint32_t circlePatchNb  = 0;
        float4 * circlePosition = nullptr; // (torus centre, circle radius)
        float4 * circleAxis     = nullptr; // (axis, probe radius)
        float4 * circleVs       = nullptr; // visibility sphere

        uint32_t concavePatchNb          = 0;
        float4 * concavePatchesPosition  = nullptr;
        int4 *   concavePatchesId        = nullptr;
        float4 * concavePatchesNeighbors = nullptr;

        float    probeRadius         = 0.f;
        uint16_t maxConcaveNeighbors = 0;
    };
} // namespace bcs::ssesdf

#endif // BCS_SSESDF_GRAPHICS_HPP
