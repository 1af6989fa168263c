/* C ABI of the MI355X solvent-excluded-surface builder (libsesdf_hip.so).
 *
 * Plain C types only; every call returns a sesdf_status and never prints. Device pointers in the
 * returned structs are owned by the handle and stay valid until the next build or destroy, as in
 * the reference (bcs::Sesdf::getData(), sesdf/src/sesdf/sesdf.cu:302-344).
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference root).
 */
#ifndef SESDF_C_H
#define SESDF_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    SESDF_OK              = 0,
    SESDF_ERR_INVALID     = 1, /* bad argument */
    SESDF_ERR_HIP         = 2, /* HIP runtime error (message in sesdf_last_error) */
    SESDF_ERR_UNSUPPORTED = 3, /* e.g. graphics = 1 (OpenGL interop is not available) */
    SESDF_ERR_CAPACITY    = 4, /* a capacity could not be grown (out of device memory) */
    SESDF_ERR_NOT_BUILT   = 5  /* data requested before the first build */
} sesdf_status;

typedef enum
{
    SESDF_MODE_COMPLETE = 0, /* bcs::Sesdf: complete surface (convex sectors, segments, concave) */
    SESDF_MODE_EXTERIOR = 1  /* bcs::Ssesdf: exterior path (torus records, concave) */
} sesdf_mode;

typedef struct sesdf_handle sesdf_handle;

/* Constructor arguments of bcs::Sesdf / bcs::Ssesdf (sesdf/include/bcs/sesdf/sesdf.hpp:18,
 * sesdf/include/bcs/ssesdf/ssesdf.hpp:17-21) plus the capacities the reference fixes at compile
 * time (sesdf/src/sesdf/sesdf.cu:15-18) and the launch environment. */
typedef struct
{
    const float * xyzr;                 /* host atoms, float4 xyz + vdW radius (non-owning, read at build) */
    uint32_t      atom_nb;
    float         probe_radius;         /* reference default 1.4 */
    float         aabb[ 6 ];            /* min xyz, max xyz of the vdW spheres; all zero: computed */
    int32_t       mode;                 /* sesdf_mode */
    uint32_t      max_neighbors;        /* per-atom neighbour capacity K (128 / 288); multiple of 32, <= 320 */
    uint32_t      max_intersection_neighbors; /* per-probe neighbour capacity (reference 32) */
    float         cell_size;            /* grid cell edge in A; 0: default */
    int32_t       device;               /* HIP device ordinal */
    void *        stream;               /* hipStream_t; NULL: the handle creates its own */
    int32_t       graphics;             /* must be 0: OpenGL interop is unsupported */
    int32_t       use_graph;            /* 1: replay the stage sequence as one hipGraph */
    int32_t       build_surface;        /* 1: build inside sesdf_create (reference buildSurface) */
} sesdf_create_info;

/* Mirror of bcs::sesdf::SesdfData (sesdf/include/bcs/sesdf/graphics.hpp:39-66). */
typedef struct
{
    uint32_t  atom_nb;
    float *   atoms;                     /* float4, grid-sorted */
    uint32_t  convex_patch_nb;
    uint32_t * convex_patches;           /* uint2 [start, end) per atom */
    uint32_t  sector_nb;
    float *   sectors;                   /* float4 (axis, half angle) */
    uint32_t  segment_patch_nb;
    uint32_t * segment_patches;          /* uint4 (i, j, intersection A, intersection B) */
    uint32_t  circle_patch_nb;
    uint32_t * circle_patches;           /* uint2 (i, j) full circles */
    uint32_t  concave_patch_nb;
    float *   concave_patches_position;  /* float4 (probe xyz, +-1 / +-2) */
    int32_t * concave_patches_id;        /* int4 (i, j, k, neighbour count) */
    float *   concave_patches_neighbors; /* float4 [max_concave_neighbors * with-neighbour count] */
    float     probe_radius;
    uint16_t  max_concave_neighbors;
} sesdf_data;

/* Mirror of bcs::ssesdf::SsesdfData (sesdf/include/bcs/ssesdf/graphics.hpp:31-50). */
typedef struct
{
    uint32_t  atom_nb;
    float *   atoms;
    uint32_t  circle_patch_nb;
    float *   circle_position;           /* float4 (torus centre, R) */
    float *   circle_axis;               /* float4 (axis, probe radius) */
    float *   circle_vs;                 /* float4 visibility sphere */
    uint32_t  concave_patch_nb;
    float *   concave_patches_position;
    int32_t * concave_patches_id;
    float *   concave_patches_neighbors;
    float     probe_radius;
    uint16_t  max_concave_neighbors;
} ssesdf_data;

/* Counts of the last build (the host counters of bcs::Sesdf, sesdf/include/bcs/sesdf/sesdf.hpp:55-80). */
typedef struct
{
    uint32_t atom_nb;
    uint32_t neighbor_nb;              /* sum of neighbour counts (ordered circles) */
    uint32_t sector_nb;                /* V */
    uint32_t intersected_circle_nb;    /* Ci */
    uint32_t full_circle_nb;           /* Cf */
    uint32_t visible_circle_nb;        /* exterior: status != 0 among i < j */
    uint32_t concave_patch_nb;         /* I */
    uint32_t concave_with_neighbors_nb;/* Iw */
    uint32_t segment_nb;               /* G */
    uint32_t torus_nb;                 /* T */
    uint32_t neighbor_overflow_atoms;  /* atoms whose neighbour list was truncated at K */
    uint32_t probe_neighbor_overflow;  /* probes with max_intersection_neighbors neighbours (truncated) */
    uint32_t circle_intersection_overflow; /* circles with more than 15 intersections */
    uint32_t jk_anomaly_nb;            /* concave records whose third circle is not intersected */
    uint32_t fp64_resolve_nb;          /* triplets re-solved in FP64 */
    uint32_t rebuild_nb;               /* capacity-growth rebuilds performed by the last build */
    uint32_t max_atom_records;
    uint64_t device_bytes;             /* device memory held by the handle */
    uint64_t build_serial;             /* stage sequences enqueued on the handle so far: changes with every build */
    uint32_t grid_relayout_nb;         /* grid lay-outs redone because atoms moved outside the cell bricks */
    uint32_t component_overflow_atoms; /* exterior SDF: atoms whose arcs exceeded the per-atom face grouping
                                          capacity (their loops were joined) */
    uint32_t max_neighbors;            /* neighbour capacity K of the last build */
    uint32_t neighbor_growth_nb;       /* K growths by the last build (SESDF_NEIGHBORS_GROW) */
} sesdf_counts;

/* Device pointers of the internal arrays exposed by the reference's analysis getters
 * (getDAtomIndices / getDAtomNeighborsCount / getDAtomNeighborsIndices / getDCircleVisibilityStatus /
 * getDGlobalToTrimmedId, sesdf/include/bcs/sesdf/sesdf.hpp:37-42). */
typedef struct
{
    uint32_t * sorted_to_initial;      /* atom_nb */
    uint32_t * neighbor_count;         /* atom_nb */
    uint32_t * neighbor_ids;           /* atom_nb * max_neighbors, sorted indices */
    uint8_t *  circle_status;          /* atom_nb * max_neighbors */
    uint32_t * trimmed_to_global;      /* intersected_circle_nb (complete surface) */
    uint32_t * circle_intersection_nb; /* intersected_circle_nb */
    uint32_t   max_neighbors;
    uint16_t   max_intersections_per_circle;
    uint16_t   max_intersection_neighbors;
} sesdf_internal;

/* bcs::Sesdf::Sesdf / bcs::Ssesdf::Ssesdf (sesdf/src/sesdf/sesdf.cu:20-75, sesdf/src/ssesdf/ssesdf.cu:19-66).
 * SESDF_ERR_INVALID, before any device work, for non-finite coordinates or radii and for a given
 * aabb that is not finite or has min > max; SESDF_ERR_CAPACITY for 2^30 atoms or more. No atoms is
 * valid (an empty surface). */
sesdf_status sesdf_create( const sesdf_create_info * info, sesdf_handle ** out );
/* bcs::Sesdf::~Sesdf (sesdf.cu:159-170). */
sesdf_status sesdf_destroy( sesdf_handle * h );

/* bcs::Sesdf::build / bcs::Ssesdf::build (sesdf.cu:172-264, ssesdf.cu:133-187): uploads the host
 * atoms given at create (or by sesdf_set_atoms) and runs the stage sequence. Blocks until counts
 * are on the host (one readback per build). */
sesdf_status sesdf_build( sesdf_handle * h );
/* Frees every device buffer of the handle (the build's neighbour lists, circles, patches, segments
 * and the sampling state) and keeps its configuration; the next build, upload or enqueue
 * allocates them again, and getters return SESDF_ERR_NOT_BUILT until then. No counterpart in the
 * reference: it serves a backward pass that rebuilds the topology instead of holding it. */
sesdf_status sesdf_release( sesdf_handle * h );
/* What a build does with an atom that has more than K neighbours. SESDF_NEIGHBORS_TRUNCATE (the
 * default): keep the K lowest sorted indices, as the reference does (operations.cu:86-121), and
 * count the atom (neighbor_overflow_atoms). SESDF_NEIGHBORS_GROW: raise K to the longest list
 * rounded up to a multiple of 32 (at most 320; beyond that lists truncate at 320) and build again.
 * No counterpart in the reference, whose K is a compile-time 128 or 288. */
#define SESDF_NEIGHBORS_TRUNCATE 0
#define SESDF_NEIGHBORS_GROW     1
sesdf_status sesdf_set_neighbor_policy( sesdf_handle * h, int policy );
/* Same stage sequence on atoms already resident in device memory (float4, atom_nb of them). When
 * d_xyzr is the handle's own input buffer (sesdf_input_buffer) nothing is copied. */
sesdf_status sesdf_build_device( sesdf_handle * h, const float * d_xyzr );
/* Replace the host atom span read by sesdf_build (same atom count). */
sesdf_status sesdf_set_atoms( sesdf_handle * h, const float * xyzr );
/* Device input buffer of the handle (float4 * atom_nb). */
sesdf_status sesdf_input_buffer( sesdf_handle * h, float ** d_xyzr );
/* Enqueue the stage sequence only (no readback, no synchronisation); counts are read later with
 * sesdf_sync. Used for timing. */
sesdf_status sesdf_enqueue( sesdf_handle * h );
sesdf_status sesdf_sync( sesdf_handle * h );

/* bcs::Sesdf::getData (sesdf.cu:302-344). */
sesdf_status sesdf_get_data( sesdf_handle * h, sesdf_data * out );
/* bcs::Ssesdf::getData (ssesdf.cu:212-232). */
sesdf_status sesdf_get_exterior_data( sesdf_handle * h, ssesdf_data * out );
sesdf_status sesdf_get_counts( sesdf_handle * h, sesdf_counts * out );
/* getMaxNeighborPerAtom ... getDGlobalToTrimmedId (sesdf.cu:346-356). */
sesdf_status sesdf_get_internal( sesdf_handle * h, sesdf_internal * out );
/* bcs::Sesdf::getAtomIndices (sesdf.cu:350): sorted -> input permutation copied to the host. */
sesdf_status sesdf_get_atom_indices( sesdf_handle * h, uint32_t * host_out );
/* Stream of the handle (hipStream_t). */
sesdf_status sesdf_get_stream( sesdf_handle * h, void ** stream );

/* -------- Stage entry points (app/detailed/sesdf_benchmark.inl:13-266 drives the reference's stages one
 * by one). Each enqueues one stage of the build on the handle's stream, on the atoms of the input
 * buffer (sesdf_upload copies the host atoms there; sesdf_input_buffer gives it for device atoms);
 * sesdf_sync ends the sequence, reads the counts back and reports SESDF_ERR_CAPACITY when a
 * capacity overflowed (sesdf_build grows capacities; these entry points do not). Stages run in
 * this order, each once per sequence: find_circles, find_intersections, build_conic_patches
 * (complete surface), handle_intersection_singularities, fill_circles (exterior path); otherwise
 * SESDF_ERR_INVALID, or SESDF_ERR_UNSUPPORTED for a stage of the other surface. -------- */
/* #1 CPU => GPU of bcs::Sesdf::build (sesdf/src/sesdf/sesdf.cu:211-213): host atoms -> input buffer. */
sesdf_status sesdf_upload( sesdf_handle * h );
/* bcs::sesdf::findCircles (sesdf/include/bcs/sesdf/operations.cuh:11-15; ssesdf operations.cuh:11):
 * grid, neighbour circles, visibility status, compaction, convex sectors. */
sesdf_status sesdf_find_circles( sesdf_handle * h );
/* bcs::sesdf::findIntersections (operations.cuh:17-18; ssesdf operations.cuh:13-14): circle-circle
 * intersections (concave patches), record placement, per-circle lists and probe positions. */
sesdf_status sesdf_find_intersections( sesdf_handle * h );
/* bcs::sesdf::buildConicPatches (operations.cuh:20-23): segments. Complete surface only. */
sesdf_status sesdf_build_conic_patches( sesdf_handle * h );
/* bcs::sesdf::handleIntersectionSingularities (operations.cuh:25-29; ssesdf operations.cuh:16-20):
 * probe neighbours of the with-neighbour concave patches. */
sesdf_status sesdf_handle_intersection_singularities( sesdf_handle * h );
/* bcs::ssesdf::fillCircles (ssesdf/include/bcs/ssesdf/operations.cuh:29-34): torus records.
 * Exterior path only. */
sesdf_status sesdf_fill_circles( sesdf_handle * h );

/* bcs::Sesdf::getDGlobalToTrimmedId (sesdf/include/bcs/sesdf/sesdf.hpp:42; filled at
 * sesdf/include/bcs/sesdf/operations.inl:46-70): for each of the atom_nb * max_neighbors + 1
 * neighbour slots, the number of intersected circles (status 1, j > i) in earlier slots. Computed on
 * request into a device array owned by the handle (valid until the next build). */
sesdf_status sesdf_get_global_to_trimmed( sesdf_handle * h, uint32_t ** d_map, uint64_t * count );

/* -------- Render-facing output: the shared-buffer layouts of bcs::Sesdf::getGraphics /
 * bcs::Ssesdf::getGraphics (sesdf/src/sesdf/sesdf.cu:266-299, sesdf/src/ssesdf/ssesdf.cu:189-210)
 * without OpenGL. The reference keeps its outputs in a few GL buffers and describes every field as
 * (buffer, byte offset, byte size); sesdf_get_graphics_layout gives the same description for the
 * last build, and sesdf_pack_graphics copies the outputs into caller-owned device buffers laid out
 * that way (for a renderer, or GL / Vulkan buffers imported into HIP). Buffers, complete surface:
 *   0 atoms [N float4]; 1 convex patches [N uint2]; 2 full circles [Cf uint2] then sectors [V float4]
 *   at Cf * 8 + (Cf * 8) % 16; 3 segments [G uint4]; 4 concave ids [I int4], positions [I float4],
 *   neighbours [Iw * max_concave_neighbors float4].
 * Exterior path: 0 atoms; 1 circle positions, axes, visibility spheres [T float4 each];
 *   2 concave positions [I float4], ids [I int4], neighbours.
 * Spans, complete: atoms, convex_patches, sectors, segment_patches, circle_patches,
 *   concave_patches_position, concave_patches_id, concave_patches_neighbors;
 * exterior: atoms, circle_position, circle_axis, circle_vs, concave_patches_position,
 *   concave_patches_id, concave_patches_neighbors (span 7 unused). -------- */
typedef struct
{
    uint32_t buffer;
    uint64_t offset;
    uint64_t size;
} sesdf_graphics_span;
typedef struct
{
    uint32_t            buffer_nb;
    uint64_t            buffer_bytes[ 5 ];
    sesdf_graphics_span spans[ 8 ];
} sesdf_graphics_layout;
sesdf_status sesdf_get_graphics_layout( sesdf_handle * h, sesdf_graphics_layout * out );
/* d_buffers: buffer_nb device pointers, buffer b holding at least buffer_bytes[b] bytes. Blocking. */
sesdf_status sesdf_pack_graphics( sesdf_handle * h, void * const * d_buffers );

/* Blocking device -> host copy (DeviceBuffer::toHost, sesdf/include/bcs/cuda/memory.cuh:155). */
sesdf_status sesdf_device_to_host( void * host, const void * device, size_t bytes );

/* -------- Molecule loading (replaces bcs::load, sesdf/src/core/loader.cpp:14-71, for MMTF files,
 * with the radius table of sesdf/src/core/molecule.cpp:9-131) -------- */
/* The atoms of the first model in file order as xyzr (host memory, 4 floats per atom, freed with
 * sesdf_free_atoms); radius from the element, upper-cased (unknown elements 1.20 A).
 * SESDF_ERR_INVALID with a message for unreadable, truncated or inconsistent files. */
sesdf_status sesdf_load_mmtf( const char * path, float ** xyzr, uint64_t * atom_nb );
void         sesdf_free_atoms( float * xyzr );
/* van der Waals radius of an element symbol (any case), as getRadius (molecule.cpp:130). */
float sesdf_element_radius( const char * element );

/* Per-stage device times of the last timed build, milliseconds (app/detailed/sesdf_benchmark.cu:50-232:
 * Circle / Intersection / Segment / Neighborhood). stage_ms receives up to SESDF_STAGE_COUNT values:
 * grid, circle kernel (neighbours + visibility), circle offsets + sectors, intersection kernel,
 * intersection offsets + placement, circle-list scan, fill, segments, probe neighbours, torus. */
#define SESDF_STAGE_COUNT 10
sesdf_status sesdf_build_timed( sesdf_handle * h, float * stage_ms, int stage_count );

/* -------- SDF sampling and its gradient (no reference counterpart: the reference ships no SDF
 * evaluator; the field is defined as rp + the signed distance to the SAS boundary of the built
 * surface, clamped to [-band, band]) -------- */
typedef struct
{
    float    origin[ 3 ];
    float    spacing[ 3 ];
    uint32_t dims[ 3 ];
} sesdf_grid_desc;

#define SESDF_SDF_EXTERIOR 1 /* flags: exterior variant (cavity surfaces excluded, cavities inside) */

/* SDF at npoints device points (float4 xyz_), written to d_out (float). Complete-surface handles.
 * From 8,192 points on, the points are evaluated in cell order (a radix sort of their cell keys in
 * a workspace the handle allocates on first use and keeps); results come back in the caller's order. */
sesdf_status sesdf_sample_points( sesdf_handle * h, const float * d_points, uint64_t npoints, float band, float * d_out );
/* Same with flags and, when d_codes is not null, the winning element of every query (uint64:
 * type << 40 | index; types 0 outside, 1 face, 2 full circle, 3 segment, 4 vertex, 5 clamped). */
sesdf_status sesdf_sample_points_ex( sesdf_handle * h, const float * d_points, uint64_t npoints, float band, int flags, float * d_out,
                                     uint64_t * d_codes );
/* Allocation-free form of sesdf_sample_points_ex. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The point sort (from 8,192 points on) runs in one device array
 * of SESDF_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on until the process
 * ends, so npoints may not exceed SESDF_NOALLOC_MAX_POINTS (SESDF_ERR_INVALID). The sampling state
 * of the current build (packed atoms and cap geometry; with SESDF_SDF_EXTERIOR the surface
 * components) must already be sized by one sesdf_sample_points_ex call with the same flags after
 * a build of this size, else SESDF_ERR_CAPACITY. Every call on one GPU uses the same array: calls
 * that overlap on a GPU (any handle, any stream) race. */
#define SESDF_NOALLOC_MAX_POINTS    1048576u
#define SESDF_NOALLOC_SCRATCH_BYTES 34079744u /* 32 bytes per point + the sort's tile histograms */
sesdf_status sesdf_sample_points_noalloc( sesdf_handle * h, const float * d_points, uint64_t npoints, float band, int flags,
                                          float * d_out, uint64_t * d_codes );
/* SDF on a regular vertex grid (x fastest), written to d_out (float, dims product). */
sesdf_status sesdf_sample_grid( sesdf_handle * h, const sesdf_grid_desc * grid, float band, float * d_out );
sesdf_status sesdf_sample_grid_ex( sesdf_handle * h, const sesdf_grid_desc * grid, float band, int flags, float * d_out );
/* d(sum_i w_i SDF(q_i)) / d(atom xyz, radius) with the patch topology of the last build held
 * fixed; d_grad receives float4 per atom in INPUT order (dx, dy, dz, dr). This form re-derives the
 * winning element of every query (recompute); sesdf_sdf_backward_ex can read stored codes. */
sesdf_status sesdf_sdf_backward( sesdf_handle * h, const float * d_points, const float * d_weights, uint64_t npoints, float band,
                                 int recompute, float * d_grad );
/* Surface components of the SAS boundary (connected pieces of faces, arcs and vertices; each bounds
 * one region of the complement of the SAS union), copied to host memory: per grid-sorted atom the
 * label of its boundary, SESDF_COMPONENT_MIXED when its exposed faces lie in several components,
 * SESDF_COMPONENT_NONE when it has no exposed boundary; and the exterior label (the component
 * holding the point of largest x of the atom with the largest x + r). */
#define SESDF_COMPONENT_MIXED 0xfffffffeu
#define SESDF_COMPONENT_NONE  0xffffffffu
sesdf_status sesdf_surface_components( sesdf_handle * h, uint32_t * host_labels, uint32_t * exterior_label );
/* d_codes: winners stored by sesdf_sample_points_ex (store) or null (recompute: from 8,192 points
 * on, a forward pass in cell order writes them to scratch codes kept with the handle, 12 bytes per
 * point); d_grad_sorted: device scratch of float4 * atom_nb (gradient in grid-sorted order). */
sesdf_status sesdf_sdf_backward_ex( sesdf_handle * h, const float * d_points, const float * d_weights, uint64_t npoints, float band,
                                    int flags, const uint64_t * d_codes, float * d_grad_sorted, float * d_grad );

/* -------- Batched build (no reference counterpart: the reference builds one molecule per
 * handle, app/benchmarks/main.cu:174-186). Molecules are concatenated molecule-major; outputs are
 * the concatenation with per-molecule offset tables. -------- */
typedef struct
{
    const float *    xyzr;          /* host atoms of all molecules, molecule-major */
    const uint64_t * atom_offsets;  /* molecule_nb + 1 */
    uint32_t         molecule_nb;
    float            probe_radius;
    int32_t          mode;
    uint32_t         max_neighbors;
    uint32_t         max_intersection_neighbors;
    float            cell_size;     /* 0: default */
    int32_t          device;
    void *           stream;
    int32_t          use_graph;
} sesdf_batch_info;

/* SESDF_ERR_INVALID, before any device work, unless atom_offsets starts at 0 and never decreases
 * (empty molecules are valid), and for non-finite coordinates or radii. */
sesdf_status sesdf_batch_create( const sesdf_batch_info * info, sesdf_handle ** out );
/* One dim^3 SDF vertex grid per molecule over its vdW bounding box grown by margin, written
 * molecule-major to d_out (molecule_nb * dim^3 floats). Complete-surface batches. */
sesdf_status sesdf_batch_sample_grids( sesdf_handle * h, uint32_t dim, float margin, float band, float * d_out );

/* Launch tuning written by the autotuner (sesdf_amd/autotune.py): grid cell edge (A, 0 keeps the
 * current one) and waves per workgroup of the circle and intersection kernels (2, 4 or 8; 0 keeps 4).
 * A new cell size re-lays out the grid; the next build recaptures the graph. */
typedef struct
{
    float    cell_size;
    uint32_t circle_waves;
    uint32_t intersection_waves;
} sesdf_tuning;
sesdf_status sesdf_set_tuning( sesdf_handle * h, const sesdf_tuning * t );
/* Per-molecule ranges into the concatenated outputs: out_offsets[(molecule_nb + 1) * 7] holds,
 * for each of atoms, sectors, full circles, concave with neighbours, concave without neighbours,
 * segments, torus records, the molecule_nb + 1 prefix offsets (host memory). */
sesdf_status sesdf_batch_offsets( sesdf_handle * h, uint64_t * out_offsets );

/* -------- Contour-Buildup comparator: bcs::ContourBuildup (sesdf/include/bcs/contour_buildup/
 * contour_buildup.hpp:24-43, sesdf/src/contour_buildup/contour_buildup.cu:58-703), the reference's
 * second surface method, which its benchmark times beside the SES builder
 * (app/benchmarks/main.cu:155-170). Same algorithm, grid (grid_dimensions^3 cells over the box,
 * reference 16) and capacities (64 neighbours per atom, 32 per probe); the OpenGL buffers become
 * device buffers owned by the handle, valid until the next build. -------- */
typedef struct sesdf_cb_handle sesdf_cb_handle;

/* ContourBuildup(molecule, aabb, probeRadius = 1.4, gridDimensions = 16, buildSurface = true)
 * (contour_buildup.hpp:27-31) plus the launch environment. */
typedef struct
{
    const float * xyzr;            /* host atoms, float4 xyz + vdW radius (non-owning, read at build) */
    uint32_t      atom_nb;         /* below 2^24 */
    float         probe_radius;
    float         aabb[ 6 ];       /* min xyz, max xyz; all zero: centre -/+ radius of the atoms (getAabb) */
    uint32_t      grid_dimensions; /* cells per axis, a power of two up to 128; 0: 16 (the reference's) */
    uint32_t      tex_size;        /* row length of the singularity texture (GL_MAX_TEXTURE_SIZE in the
                                      reference, contour_buildup.cu:182-186); 0: 32768 */
    int32_t       device;
    void *        stream;          /* hipStream_t; NULL: the handle creates its own */
    int32_t       use_graph;       /* 1: replay the stage sequence as one hipGraph */
    int32_t       build_surface;   /* 1: build inside sesdf_cb_create */
} sesdf_cb_create_info;

/* The buffers behind ContourBuildupGraphics (sesdf/include/bcs/contour_buildup/graphics.hpp:6-19;
 * contour_buildup.cu:97-112 buffer roles). Device pointers, float4 unless noted. */
typedef struct
{
    uint32_t atom_nb;
    float *  atoms;                    /* grid-sorted atoms */
    uint32_t probe_nb;                 /* spherical triangles */
    float *  probe_positions;          /* PROBE_POS: xyz, probe radius */
    float *  sphere_triangle_vec1;     /* SPHERE_TRIA_VEC_1: atom i - probe, 1 */
    float *  sphere_triangle_vec2;     /* SPHERE_TRIA_VEC_2: atom j - probe, 1 */
    float *  sphere_triangle_vec3;     /* SPHERE_TRIA_VEC_3: atom k - probe, probe radius^2 */
    uint32_t torus_nb;
    float *  torus_positions;          /* TORUS_POS: centre, radius R */
    float *  torus_axes;               /* TORUS_AXIS: unit axis, probe radius */
    float *  torus_visibility_spheres; /* TORUS_VS: centre (relative to the torus centre), radius */
    float *  tex_coords;               /* TEX_COORD: float3 per probe (neighbour count, row, column) */
    float *  singularity_texture;      /* SING_TEX: float3 * probe_neighbor_nb per sorted probe (first
                                          count entries set) */
    uint32_t probe_neighbor_nb;        /* 32 */
    uint32_t tex_size;
    float    probe_radius;
} sesdf_cb_data;

typedef struct
{
    uint32_t atom_nb;
    uint32_t probe_nb;                 /* numProbes (contour_buildup.cu:467-481) */
    uint32_t torus_nb;                 /* numSC (contour_buildup.cu:452-463) */
    uint32_t neighbor_overflow_atoms;  /* atoms with more than 64 neighbours (lists truncated, as the reference) */
    uint32_t no_contour_atoms;         /* atoms whose contour vanished in the covered-circle pass */
    uint32_t arc_capacity;             /* arcs per circle held on chip (8, or 80 after an overflow) */
    uint32_t build_attempts;           /* builds run by the last call (capacity growth repeats it) */
    uint32_t grid_dimensions;
    uint32_t singularity_grid;         /* 1: the singularity texture came from the 2 rp search grid */
    uint64_t device_bytes;
    uint64_t build_serial;
} sesdf_cb_counts;

/* Internal arrays (device pointers), for census and parity checks. */
typedef struct
{
    uint32_t * sorted_to_initial; /* N: sorted atom -> input atom */
    uint32_t * neighbor_count;    /* N: neighbours after the covered-circle pass (0: no contour) */
    uint32_t * neighbor_ids;      /* N * max_neighbors: sorted atom indices, scan order */
    uint32_t   max_neighbors;     /* 64 */
    uint32_t * probe_atoms;       /* uint4 per probe: (i, j, k, 0) sorted atom indices */
    uint32_t * torus_atoms;       /* uint2 per torus: (i, j) */
    uint32_t * probe_sorted;      /* per sorted probe: probe index (row order of the texture) */
} sesdf_cb_internal;

/* ContourBuildup::ContourBuildup (contour_buildup.cu:58-369). */
sesdf_status sesdf_cb_create( const sesdf_cb_create_info * info, sesdf_cb_handle ** out );
/* ContourBuildup::~ContourBuildup (contour_buildup.cu:370-398). */
sesdf_status sesdf_cb_destroy( sesdf_cb_handle * h );
/* ContourBuildup::build (contour_buildup.cu:400-660): uploads the host atoms, builds, reads the
 * counts back (blocking). */
sesdf_status sesdf_cb_build( sesdf_cb_handle * h );
/* Same on atoms already in device memory (float4, atom_nb of them; null: the input buffer). */
sesdf_status sesdf_cb_build_device( sesdf_cb_handle * h, const float * d_xyzr );
/* Replace the host atom span read by sesdf_cb_build (same atom count). */
sesdf_status sesdf_cb_set_atoms( sesdf_cb_handle * h, const float * xyzr );
sesdf_status sesdf_cb_input_buffer( sesdf_cb_handle * h, float ** d_xyzr );
/* Replay the last build's stage sequence without readback (timing); sesdf_cb_sync ends it and
 * reports SESDF_ERR_CAPACITY if a replay outgrew a capacity. */
sesdf_status sesdf_cb_enqueue( sesdf_cb_handle * h );
sesdf_status sesdf_cb_sync( sesdf_cb_handle * h );
sesdf_status sesdf_cb_get_stream( sesdf_cb_handle * h, void ** stream );
/* ContourBuildup::getGraphics (contour_buildup.cu:662-665): the buffers and counts. */
sesdf_status sesdf_cb_get_data( sesdf_cb_handle * h, sesdf_cb_data * out );
sesdf_status sesdf_cb_get_counts( sesdf_cb_handle * h, sesdf_cb_counts * out );
sesdf_status sesdf_cb_get_internal( sesdf_cb_handle * h, sesdf_cb_internal * out );
/* Capacities and search of the next build (no reference counterpart: the reference sizes its
 * buffers for 64 probes per atom). arc_capacity: arcs per circle held on chip, 8 (default; wider
 * lists build again at 80) or 80; probe_capacity / torus_capacity: output slots (0 keeps the
 * current ones; a build that needs more grows them and runs again); singularity_search: how the
 * singularity texture finds probe neighbours, 0 = by cell size (default), 1 = the reference's
 * stencil scan, 2 = a grid of 2 rp cells (same texture). */
typedef struct
{
    uint32_t arc_capacity;
    uint32_t probe_capacity;
    uint32_t torus_capacity;
    uint32_t singularity_search;
} sesdf_cb_tuning;
sesdf_status sesdf_cb_set_tuning( sesdf_cb_handle * h, const sesdf_cb_tuning * t );

const char * sesdf_status_string( sesdf_status s );
const char * sesdf_last_error( void );

#ifdef __cplusplus
}
#endif

#endif /* SESDF_C_H */
