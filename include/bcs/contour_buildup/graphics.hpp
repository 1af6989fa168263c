// bcs::cb::ContourBuildupGraphics: what the reference's Contour-Buildup hands its renderer
// (sesdf/include/bcs/contour_buildup/graphics.hpp:6-19): counts and OpenGL vertex-array / texture
// names. There is no GL interop here, so the names are 0; ContourBuildupData carries the device
// buffers those vertex arrays would bind (sesdf/src/contour_buildup/contour_buildup.cu:97-112).
// Device pointers use HIP's vector types; compile with hipcc, or with a host compiler and
// -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include.
#ifndef BCS_CONTOURBUILDUP_GRAPHICS_HPP
#define BCS_CONTOURBUILDUP_GRAPHICS_HPP

#include <cstddef>
#include <cstdint>

#include <hip/hip_vector_types.h>

#include "bcs/core/type.hpp"

using GLuint = unsigned int;

namespace bcs::cb
{
    struct ContourBuildupGraphics
    {
        std::size_t atomNb               = 0;
        GLuint      atomVao              = 0;
        std::size_t sphericalTriangleNb  = 0;
        GLuint      sphericalTriangleVao = 0;
        std::size_t torusNb              = 0;
        GLuint      torusVao             = 0;
        GLuint      singularityTexture   = 0;
    };

    // Device buffers of the last build (owned by the ContourBuildup object, valid until its next
    // build). float4 unless noted.
    struct ContourBuildupData
    {
        uint32_t atomNb = 0;
        float4 * atoms  = nullptr;            // grid-sorted atoms
        uint32_t sphericalTriangleNb = 0;
        float4 * probePositions      = nullptr; // xyz, probe radius
        float4 * sphereTriangleVec1  = nullptr; // atom i - probe, 1
        float4 * sphereTriangleVec2  = nullptr; // atom j - probe, 1
        float4 * sphereTriangleVec3  = nullptr; // atom k - probe, probe radius^2
        uint32_t torusNb             = 0;
        float4 * torusPositions      = nullptr; // centre, radius R
        float4 * torusAxes           = nullptr; // unit axis, probe radius
        float4 * torusVisibilitySpheres = nullptr;
        float *  texCoords           = nullptr; // float3 per probe: neighbour count, row, column
        float *  singularityTexture  = nullptr; // float3 * 32 per sorted probe
        uint32_t probeNeighborNb     = 32;
        uint32_t texSize             = 0;
        float    probeRadius         = 0.f;
    };
} // namespace bcs::cb

#endif // BCS_CONTOURBUILDUP_GRAPHICS_HPP
