/*
 * slr_hip.h - C ABI of the MI355X spectral path tracer (libslr_hip.so).
 *
 * The library is a drop-in for the hot path of SLR: the scene it renders,
 * the SBVH-collapsed-to-QBVH accelerator, the unidirectional spectral path
 * tracer with MIS, and the film. Every entry point returns an int status
 * (SLRH_OK or a negative error code); no C++ exception crosses the ABI;
 * slrh_last_error() describes the last failure on a context. Host buffers
 * are owned by the caller, device buffers by the library unless an entry
 * point says it takes device pointers. Calls on the scenes and films of a
 * context may come from several threads: they serialise on the context's
 * lock (a render's snapshot callback runs with it held by the rendering
 * thread), and slrh_last_error() returns the calling thread's last failure
 * on that context.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the SLR checkout).
 */
#ifndef SLR_HIP_H
#define SLR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLRH_API __attribute__((visibility("default")))

enum slrh_status {
    SLRH_OK = 0,
    SLRH_ERR_INVALID = -1,   /* bad argument */
    SLRH_ERR_STATE = -2,     /* call out of order (e.g. query before build) */
    SLRH_ERR_DEVICE = -3,    /* HIP runtime error */
    SLRH_ERR_NOMEM = -4,     /* device or host allocation failed */
    SLRH_ERR_CAPACITY = -5,  /* request exceeds a documented limit */
    SLRH_ERR_NOTFOUND = -6   /* unknown ticket / id */
};

typedef struct slrh_context slrh_context;
typedef struct slrh_scene slrh_scene;
typedef struct slrh_film slrh_film;
typedef struct slrh_server slrh_server;

/* ------------------------------------------------------------ lifecycle */

/* Creates a context on HIP device `device` (device < 0: a host-only context
 * that assembles scenes and builds/inspects the accelerator but cannot
 * trace or render). Replaces the process-wide state
 * set up by SLR::initializeColorSystem() (libSLR/BasicTypes/spectrum_base.cpp:224)
 * and the per-render thread pool (libSLR/Helper/ThreadPool.h:17-84). */
SLRH_API int slrh_create(int device, slrh_context** out);
SLRH_API int slrh_destroy(slrh_context* ctx);
SLRH_API const char* slrh_last_error(const slrh_context* ctx);
SLRH_API int slrh_version(void);

/* ------------------------------------------------------------ spectra */

enum slrh_spectrum_type {
    SLRH_SPECTRUM_REFLECTANCE = 0,  /* SpectrumType::Reflectance */
    SLRH_SPECTRUM_LIGHT_SOURCE = 1  /* SpectrumType::LightSource */
};

/* RGB (linear sRGB) upsampled to a spectrum by Meng et al. 2015; replaces
 * UpsampledContinuousSpectrum(spType, ColorSpace::sRGB, r, g, b)
 * (libSLR/BasicTypes/spectrum_types.cpp:482-543). */
SLRH_API int slrh_spectrum_rgb(slrh_scene* s, int type, float r, float g, float b, int* id);
/* Regularly sampled spectrum over [lambda_min, lambda_max] in nm; replaces
 * RegularContinuousSpectrum (libSLR/BasicTypes/spectrum_types.h:88-108).
 * n == 1 gives a constant. */
SLRH_API int slrh_spectrum_regular(slrh_scene* s, float lambda_min, float lambda_max,
                                   const float* values, int n, int* id);

/* Evaluates spectrum `id` at n wavelengths (nm) on the host, as the device
 * does during rendering (RegularContinuousSpectrum::evaluate,
 * libSLR/BasicTypes/spectrum_types.cpp:82-100). */
SLRH_API int slrh_spectrum_evaluate(const slrh_scene* s, int id, int n, const float* lambdas, float* out);
/* The film's colour-matching strata and their luminance integral
 * (DiscretizedSpectrum::initialize, libSLR/BasicTypes/spectrum_types.cpp:945-990). */
SLRH_API int slrh_cmf_strata(float* xbar16, float* ybar16, float* zbar16, float* integral);

/* ------------------------------------------------------------ materials */

/* DiffuseReflectionSurfaceMaterial without sigma -> LambertianBRDF
 * (libSLR/SurfaceMaterial/basic_surface_materials.cpp:16-26). */
SLRH_API int slrh_material_lambert(slrh_scene* s, int reflectance, int* id);
/* DiffuseReflectionSurfaceMaterial with sigma -> OrenNayerBRDF
 * (libSLR/BSDF/OrenNayerBRDF.h:28-30). */
SLRH_API int slrh_material_oren_nayar(slrh_scene* s, int reflectance, float sigma, int* id);
/* MicrofacetReflectionSurfaceMaterial, GGX (libSLR/SurfaceMaterial/microfacet_surface_materials.cpp:15-20). */
SLRH_API int slrh_material_ggx_conductor(slrh_scene* s, int eta, int k, float alpha_x, float alpha_y, int* id);
/* MicrofacetScatteringSurfaceMaterial, GGX (microfacet_surface_materials.cpp:24-30). */
SLRH_API int slrh_material_ggx_dielectric(slrh_scene* s, int eta_ext, int eta_int, float alpha_x,
                                          float alpha_y, int* id);
/* EmitterSurfaceMaterial(base, DiffuseEmitterSurfaceProperty(emittance))
 * (libSLR/Core/surface_material.h:55-70). base may be -1 only if no path
 * ever needs to scatter off the emitter (the reference then dereferences null). */
SLRH_API int slrh_material_diffuse_emitter(slrh_scene* s, int base_material, int emittance, int* id);

/* ------------------------------------------------------------ geometry */

/* Adds a triangle mesh in world space. Replaces SLR::TriangleMeshNode +
 * MaterialGroupInTriangleMesh::setTriangles (libSLR/Scene/TriangleMeshNode.h:17-35)
 * with its Vertex{position, normal, tangent, texCoord} (libSLR/Core/geometry.h:293-301).
 * pos: 3*nverts floats; nrm/tan: 3*nverts floats or NULL (then the geometric
 * normal / an orthogonal tangent is used); idx: 3*ntris; mat: ntris material
 * ids, or NULL with `material` used for every triangle. */
SLRH_API int slrh_mesh_add(slrh_scene* s, uint32_t nverts, const float* pos, const float* nrm,
                           const float* tan, uint32_t ntris, const uint32_t* idx,
                           const uint32_t* mat, int material, int* mesh_id);

/* Adds a mesh that appears only through instances (a prototype), with the
 * arguments of slrh_mesh_add, in object space. Replaces a TriangleMeshNode
 * wrapped in a ReferenceNode (libSLR/Scene/node.h:68-82, node.cpp ReferenceNode::
 * createRenderingData), which builds the mesh's own aggregate once and shares
 * it between every place it is referenced from. */
SLRH_API int slrh_prototype_add(slrh_scene* s, uint32_t nverts, const float* pos, const float* nrm,
                                const float* tan, uint32_t ntris, const uint32_t* idx,
                                const uint32_t* mat, int material, int* proto_id);
/* Places a copy of prototype proto_id with the object-to-world affine
 * transform l2w (12 floats: 3 rows of 4, row-major; the bottom row is
 * 0 0 0 1). material >= 0 replaces every material of the prototype for this
 * copy, -1 keeps them (an extension: the reference binds materials to the
 * mesh). Replaces InternalNode(StaticTransform) over a ReferenceNode, which
 * becomes a TransformedSurfaceObject (libSLR/Scene/node.cpp InternalNode::
 * createRenderingData; libSLR/Core/surface_object.cpp:290-369): rays are
 * taken into object space by the inverse transform, hits and light samples
 * back by the transform. Instance ids count from 0 in call order. */
SLRH_API int slrh_instance_add(slrh_scene* s, int proto_id, const float* l2w, int material, int* inst_id);

/* The transform of instance inst_id as stored (either pointer may be NULL):
 * l2w, the 12 object-to-world floats given (row-major 3x4), and w2l, the 16
 * world-to-object floats (row-major 4x4) rays are transformed with,
 * computed as the reference inverts a StaticTransform's matrix (Matrix4x4
 * invert, libSLR/BasicTypes/Matrix4x4.cpp:36-87). Its bottom row carries the
 * elimination's FP32 rounding residue; a point is divided by the w it gives
 * when w != 1 (Matrix4x4 * Point3D, Matrix4x4.h:68-74). */
SLRH_API int slrh_instance_get(const slrh_scene* s, int inst_id, float* l2w, float* w2l);

enum slrh_compat_flags {
    /* shadow rays test an instance with the untransformed world ray, as
     * TransformedSurfaceObject::testVisibility does (libSLR/Core/
     * surface_object.cpp:363-369); off: the ray is transformed as for
     * closest hits */
    SLRH_COMPAT_INSTANCE_SHADOW = 1,
    /* SLRH_VOL_PT renders treat a light sample in the plane of the shading
     * surface as the reference VolumetricPTRenderer does: its next-event
     * estimate has no cosine guard (libSLR/Renderer/VolumetricPTRenderer.cpp:
     * 190-231), the BSDF's shading-normal correction is then 0/0
     * (libSLR/Core/directional_distribution_functions.h:325-327), and the NaN
     * drops the whole sample (VolumetricPTRenderer.cpp:133-138). Such samples
     * are counted in the non-finite statistic. Off: the estimate is zero, as
     * in SLRH_PT (PTRenderer.cpp:199), and a scene without media renders
     * exactly as with SLRH_PT. */
    SLRH_COMPAT_VOLPT_TANGENT_NEE = 2
};
/* A heterogeneous medium on a density lattice: region bounds[6] (min xyz,
 * max xyz), nx*ny*nz densities (x fastest, then y, then z; trilinear
 * between lattice points spanning the region), base scattering and
 * extinction spectra (the extinction constant over wavelength) and the
 * Henyey-Greenstein asymmetry g. Replaces DensityGridMediumNode(region,
 * base_sigma_s, base_sigma_e, density_grid, nx, ny, nz,
 * HenyeyGreensteinScatteringMediumMaterial(g)) (libSLR/Scene/medium_nodes.h:
 * 33-46). Free paths are sampled against a majorant grid of the maximum
 * density over blocks of 8^3 lattice cells (the reference uses trilinear
 * supervoxels with a per-cell correction, DensityGridMediumDistribution.cpp:
 * 45-205; both bound the density, so both are unbiased). One medium per
 * scene; it takes part in SLRH_VOL_PT renders. */
SLRH_API int slrh_medium_density_grid(slrh_scene* s, const float* bounds, const float* density, uint32_t nx,
                                      uint32_t ny, uint32_t nz, int sigma_s, int sigma_e, float g,
                                      int* medium_id);

/* Reproduces reference behaviours that differ from the correct result, for
 * parity checks (flags from slrh_compat_flags). Call before slrh_scene_build. */
SLRH_API int slrh_scene_compat(slrh_scene* s, uint32_t flags);

/* Camera-to-world 4x4 (column-major) and PerspectiveCamera parameters
 * (libSLR/Camera/PerspectiveCamera.cpp:16-27, libSLR/Scene/camera_nodes.h:16-35). */
/* Image-based environment light: an equirectangular image of width*height
 * linear sRGB texels (row 0 at the +y pole, phi = 0 towards +z, as
 * Vector3D::toPolarYUp maps directions), radiance scale. Replaces
 * InfiniteSphereNode(scene, ImageSpectrumTexture, scale) set with
 * Scene::setEnvironmentNode (libSLR/Scene/node.h:133, Scene.h:33); the
 * texels are kept as the reference keeps an RGBA16F light-source image
 * (uvs truncated to half precision, libSLR/Core/image_2d.h:270-289). Call
 * before slrh_scene_build. */
SLRH_API int slrh_environment(slrh_scene* s, const float* rgb, uint32_t width, uint32_t height, float scale);

SLRH_API int slrh_camera_perspective(slrh_scene* s, const float* l2w, float sensitivity, float aspect,
                                     float fov_y, float lens_radius, float img_dist, float obj_dist);

/* Scene file reader: runs the text of a scene written in SLR's scene
 * language (libSLRSceneGraph/Parser/SceneParser.yy, the builtins of
 * libSLRSceneGraph/API.cpp:295-1325; readScene in HostProgram/main.cpp:40)
 * and adds what it describes to `s`: spectra (RGB, regular, irregular —
 * tabulated over its knots — and the illuminants D65 and E), materials
 * ("matte" with or without sigma, "microfacet metal", "microfacet glass",
 * "emitter" with a "diffuse" emitter property), meshes in world space with
 * the vertex normals and tangents the reference recomputes, and the
 * perspective camera. Control flow, user functions, model and image loading
 * and other materials and renderers are refused with a message naming the
 * line. `info` (may be NULL) receives the render settings the file sets. */
typedef struct slrh_scene_file_info {
    uint32_t width, height;   /* setRenderSettings (1024 x 1024 when not set) */
    uint32_t spp;             /* setRenderer's "samples" (8 when not set) */
    uint32_t integrator;      /* slrh_integrator: "PT", "Volumetric PT" or "BPT" */
    uint32_t rng_seed;        /* setRenderSettings "rngSeed" */
    float brightness;         /* setRenderSettings "brightness": saveImage's scale is brightness / spp */
    float sensitivity;        /* the camera's image sensor: its sensitivity, or 1 / (pi r^2) when that is
                                 not positive (PerspectiveCamera.cpp:24; infinite for a pinhole) */
    uint32_t num_meshes, num_triangles, num_materials, num_spectra;
} slrh_scene_file_info;
SLRH_API int slrh_scene_load_text(slrh_scene* s, const char* text, slrh_scene_file_info* info);

/* Introspection (no reference counterpart; for tests and integrators that
 * check what a scene holds). slrh_scene_geometry copies the world-level
 * triangles: on entry *nverts / *ntris are the capacities of the arrays
 * (arrays may be NULL), on return the counts; pos/nrm/tan 3 floats per
 * vertex, idx 3 per triangle, mat 1 per triangle. slrh_scene_camera_get:
 * l2w (16 floats, column-major) and params {sensitivity, aspect, fovY,
 * lens radius, image distance, object distance}. slrh_material_info: the
 * material's model (0 Lambert, 1 Oren-Nayar, 2 GGX conductor, 3 GGX
 * dielectric, 4 emitter without a base), its spectra (-1: none), emittance
 * spectrum and two parameters (Oren-Nayar A, B; GGX alpha_x, alpha_y). */
SLRH_API int slrh_scene_geometry(const slrh_scene* s, uint32_t* nverts, uint32_t* ntris, float* pos, float* nrm,
                                 float* tan, uint32_t* idx, uint32_t* mat);
SLRH_API int slrh_scene_camera_get(const slrh_scene* s, float* l2w, float* params);
SLRH_API int slrh_material_info(const slrh_scene* s, int id, int* model, int* spec0, int* spec1, int* emittance,
                                float* p0, float* p1);

/* Builds the SBVH (libSLR/Accelerator/SBVH.h:380-417), collapses it into the
 * QBVH (libSLR/Accelerator/QBVH.h:252-284) and uploads scene, accelerator and
 * tables to HBM. Replaces SLR::Scene::build (libSLR/Scene/Scene.cpp:17-38). */
SLRH_API int slrh_scene_create(slrh_context* ctx, slrh_scene** out);
SLRH_API int slrh_scene_build(slrh_scene* s);
SLRH_API int slrh_scene_destroy(slrh_scene* s);

/* Accelerators built besides the SBVH collapsed to the QBVH (which renders):
 * SLRH_ACCEL_STANDARD_BVH adds the binned-SAH binary BVH the reference builds
 * as StandardBVH(objs, Partitioning::BinnedSAH) (libSLR/Accelerator/
 * StandardBVH.h:62-193, 16 bins, the same tree node for node) for ray-batch
 * queries. Scenes without instances only. Call before slrh_scene_build. */
enum slrh_accel_flags { SLRH_ACCEL_STANDARD_BVH = 1 };
SLRH_API int slrh_scene_accelerators(slrh_scene* s, uint32_t flags);
/* Which accelerator slrh_intersect / slrh_occluded traverse: the QBVH
 * (QBVH::intersect, the default) or the binned-SAH binary BVH
 * (StandardBVH::intersect / testVisibility, StandardBVH.h:265-330). Hits are
 * the same up to ties at equal t. */
enum slrh_query_accel { SLRH_QUERY_QBVH = 0, SLRH_QUERY_STANDARD_BVH = 1 };
SLRH_API int slrh_scene_query_accel(slrh_scene* s, int accel);
/* Dump of the binned-SAH BVH in the layout of slrh_scene_sbvh_dump (nodes in
 * StandardBVH's m_nodes order); stats[0] depth, stats[1] SAH cost
 * (StandardBVH::calcSAHCost). Any output may be NULL. */
SLRH_API int slrh_scene_standard_bvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs,
                                          uint32_t* num_nodes, uint32_t* num_refs, double* stats);

/* stats[0..7]: SBVH nodes, SBVH leaf refs, SBVH depth, SBVH SAH cost,
 * QBVH nodes, QBVH leaf refs, QBVH depth, QBVH SAH cost; stats[8]: build seconds. */
SLRH_API int slrh_scene_accel_stats(const slrh_scene* s, double* stats);
/* stats[0]: seconds of all the scene's SBVH builds and QBVH collapses (world
 * tree and prototype trees); stats[1]: SBVH nodes of them binned and
 * partitioned on the device. With a device context, the nodes of at least
 * 2^16 fragments at the top of a triangle tree keep their fragments in
 * device memory and are binned and partitioned there, with the same
 * results as the host builder (the tree is the same node for node); the
 * host builds the rest of the tree on its threads. */
SLRH_API int slrh_scene_build_stats(const slrh_scene* s, double* stats);
/* Topology dumps in the same layout as the reference's node arrays (for
 * parity tests): SBVH boxes[6*n], info[5*n] {c0,c1,offset,numLeaves,axis|0xFF},
 * refs[leaf refs]; QBVH boxes[24*n], info[7*n] {children[4] in the reference's
 * bitfield encoding, top/left/right axis}, refs. Buffers may be NULL to query n.
 * These are the world-level trees; a ref is a global triangle index, or
 * INT32_MIN + i (bit 31 set) for instance i. */
SLRH_API int slrh_scene_sbvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs,
                                  uint32_t* num_nodes, uint32_t* num_refs);
SLRH_API int slrh_scene_qbvh_dump(const slrh_scene* s, float* boxes, uint32_t* info, int32_t* refs,
                                  uint32_t* num_nodes, uint32_t* num_refs);

/* ------------------------------------------------------------ ray queries */

/* Structure-of-arrays ray batch; all pointers refer to n-element arrays. */
typedef struct slrh_rays {
    const float* ox; const float* oy; const float* oz;
    const float* dx; const float* dy; const float* dz;
    const float* tmin; const float* tmax;
} slrh_rays;

typedef struct slrh_hits {
    float* t;        /* hit distance, or tmax on a miss */
    int32_t* prim;   /* global triangle index (meshes and prototypes in the order they were
                        added, then triangle order), -1 on a miss */
    float* b0;       /* surface parameters as SurfaceInteraction stores them: */
    float* b1;       /*   b0 = 1 - u - v (weight of vertex 0), b1 = u (weight of vertex 1) */
    int32_t* inst;   /* instance of the hit triangle, -1 for world-space triangles and
                        misses; may be NULL */
} slrh_hits;

enum slrh_mem { SLRH_MEM_HOST = 0, SLRH_MEM_DEVICE = 1 };

/* Closest hit for n rays. Replaces QBVH::intersect (libSLR/Accelerator/QBVH.h:355-365)
 * over TriangleSurfaceShape::intersect (libSLR/SurfaceShape/TriangleSurfaceShape.cpp:205-257).
 * mem says whether rays/hits are host or device pointers. stream: a hipStream_t or NULL. */
SLRH_API int slrh_intersect(slrh_scene* s, const slrh_rays* rays, const slrh_hits* hits, uint32_t n,
                            int mem, void* stream);
/* Visibility (1 unoccluded, 0 occluded) for n segments [tmin, tmax]. Replaces
 * QBVH::testVisibility (QBVH.h:367-376) over TriangleSurfaceShape::testVisibility
 * (TriangleSurfaceShape.cpp:259-299). */
SLRH_API int slrh_occluded(slrh_scene* s, const slrh_rays* rays, float* visibility, uint32_t n,
                           int mem, void* stream);

/* Traversal work of the scene's last slrh_intersect / slrh_occluded batch
 * through the QBVH (no reference counterpart; for measurement): QBVH nodes
 * popped (each tests its four child boxes) and triangles tested, summed over
 * the batch; the counts an instrumented QBVH::commonProcedure
 * (libSLR/Accelerator/QBVH.h:294-341) gives for the same rays. */
SLRH_API int slrh_scene_query_counts(const slrh_scene* s, double* nodes, double* tris);

/* slrh_intersect for device pointers that allocates nothing: no counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). rays and hits are device pointers; hits as slrh_intersect
 * writes them. The traversal's counters and stack spill live in a static
 * device array of 48 MiB that each GPU the library runs on holds until the
 * process ends (looked up when a context is created on that device); its
 * grid is at most 1024 blocks. The call is asynchronous on `stream` (a
 * hipStream_t, NULL for the default stream) and reads nothing back: results
 * are valid once the stream has completed. Calls that overlap on one GPU
 * share that array and race: serialise them (one stream, or one at a time).
 * Returns SLRH_ERR_CAPACITY for n > 2^30 and for a scene whose trees could
 * need a traversal stack deeper than the reference's 64 entries. */
SLRH_API int slrh_intersect_noalloc(slrh_scene* s, const slrh_rays* rays, const slrh_hits* hits, uint32_t n,
                                    void* stream);

enum slrh_medium_query_mode {
    SLRH_MEDIUM_TRANSMITTANCE = 0,  /* one ratio-tracking estimate of the segment's transmittance */
    SLRH_MEDIUM_FREE_PATH = 1       /* one delta-tracking scattering distance, tmax if none */
};
/* Ray batches through the scene's medium (surfaces are not tested): per ray
 * an unbiased transmittance estimate of [tmin, tmax], as Scene::testVisibility
 * takes from MediumObjectAggregate::evaluateTransmittance (libSLR/Core/
 * medium_object.cpp:348-399), or a sampled free-path distance, as
 * Scene::interact takes from MediumObjectAggregate::interact (:300-347).
 * Ray i uses its own random stream seeded from (seed, i). out: n floats. */
SLRH_API int slrh_medium_query(slrh_scene* s, const slrh_rays* rays, uint32_t n, int mode, uint32_t seed,
                               float* out, int mem, void* stream);

/* ------------------------------------------------------------ rendering */

/* SLRH_PT: PTRenderer (libSLR/Renderer/PTRenderer.cpp), media ignored.
 * SLRH_VOL_PT: VolumetricPTRenderer (libSLR/Renderer/VolumetricPTRenderer.cpp:
 * 146-296): free paths sampled by delta tracking in the scene's medium,
 * scattering by its phase function, light samples attenuated by the
 * medium's transmittance (ratio tracking), no fixed path-length cap (Russian
 * roulette; max_depth still applies when set). */
/* SLRH_BPT: BPTRenderer (libSLR/Renderer/BPTRenderer.cpp:25-432): per pixel
 * sample a light subpath from a light chosen as Scene::selectSurfaceLight
 * does and an eye subpath, both ended by Russian roulette, and every
 * connection (s, t) between them, weighted by the reference's power-heuristic
 * MIS over subpath area PDFs and RR probabilities; light subpath vertices
 * connected to the lens add to the pixel their direction crosses. Media are
 * ignored (as in BPTRenderer); a scene with an environment light is refused.
 * max_depth caps the eye subpath's vertices (0: none; the reference has no
 * cap); at most 64 light subpath vertices are kept per sample (stats[22]
 * counts subpaths stopped by either cap). stats[16] light and [17] eye
 * subpath vertices, [19] connections traced, [20] lens contributions filed.
 * Each of the pixels x slots_per_pixel slots (at most 2^24; automatic:
 * about 4M, 16 KB of device memory each) renders its pixel's samples one
 * after another from its own random stream, at most 65535 per slot; films
 * of equal seeds and slots_per_pixel are the same however the device
 * interleaves the slots. The lens
 * contributions are summed with FP64 atomics in a varying order, so
 * films of equal seeds agree to rounding (about 1e-15 relative), not
 * bit-for-bit. slrh_render_batch renders BPT frames one after another. */
enum slrh_integrator { SLRH_PT = 0, SLRH_VOL_PT = 1, SLRH_BPT = 2 };

enum slrh_render_flags {
    SLRH_FLAG_TIME_KERNELS = 1,   /* time the kernel groups with HIP events (stats 5, 8, 9) */
    SLRH_FLAG_NO_GRAPH = 2,       /* launch each iteration's kernels instead of replaying a recorded graph */
    SLRH_FLAG_PARALLEL_SHADE = 4, /* run the conductor, dielectric and miss shading kernels on forked
                                     streams beside the diffuse one (graph branches); off by default:
                                     replayed graphs with these branches measured slower than in order */
    SLRH_FLAG_SHADE_BLOCK_128 = 8, /* surface shading kernels in 128-thread workgroups; without it they
                                     use the autotuner's choice (tune_trace 0) or 256 */
    SLRH_FLAG_TUNE_POOL = 16      /* with slots_per_pixel 0: the wavefront width comes from a pilot
                                     (first such render of the scene) that renders two samples per
                                     slot at 8M, 16M and 32M live paths and keeps the most samples
                                     per second; without it about 32M. Films depend on the slots per
                                     pixel, so with this flag they depend on the pilot's timing */
};

typedef struct slrh_render_settings {
    uint32_t width, height;   /* RenderSettingItem::ImageWidth/ImageHeight */
    uint32_t spp;             /* PTRenderer(spp) */
    uint32_t max_depth;       /* 0 = the reference's cap of 100 bounces; else a tighter cap */
    uint32_t seed;            /* RenderSettingItem::RNGSeed */
    uint32_t integrator;      /* slrh_integrator */
    uint32_t slots_per_pixel; /* live paths per pixel (0 = automatic) */
    uint32_t flags;           /* SLRH_FLAG_* */
    uint32_t tune_trace;      /* traversal kernel variant + 1; 0 = automatic: the first render
                                 of a scene with a pool of >= 2^20 slots renders one pilot
                                 sample per slot with each variant (scratch film) and keeps
                                 the fastest per traced ray, then pilots the shading
                                 kernels' workgroup sizes (256, 128) with that variant and
                                 keeps the faster per shaded vertex, for every later render
                                 of the scene (variant 0 and 256 until then); neither
                                 choice changes the film */
    uint32_t reserved;        /* must be 0 */
} slrh_render_settings;

/* Traversal kernel variants. On entry *n is the capacity of the arrays (any
 * may be NULL), on return the number of variants; for variant v: its LDS
 * stack depth per lane, persistent-refill threshold (idle lanes) and whether
 * a node that opens a leaf has the leaf's first triangle tested in the same
 * loop iteration. */
SLRH_API int slrh_trace_variants(int* n, int* lds_depth, int* refill, int* fused);

/* Film: FP64 16-stratum spectral sums per pixel, the same quantity the
 * reference's ImageSensor holds (libSLR/Core/ImageSensor.cpp:125-130,
 * libSLR/BasicTypes/spectrum_types.h:716-741) but summed in FP64. */
SLRH_API int slrh_film_create(slrh_context* ctx, uint32_t width, uint32_t height, slrh_film** out);
SLRH_API int slrh_film_destroy(slrh_film* f);
SLRH_API int slrh_film_clear(slrh_film* f);

enum slrh_film_format {
    SLRH_FILM_STRATA16_F64 = 0,  /* h*w*16 doubles: raw stratum sums */
    SLRH_FILM_XYZ_F64 = 1,       /* h*w*3 doubles: CIE XYZ of the sums (DiscretizedSpectrum::getRGB's XYZ step) */
    SLRH_FILM_SRGB8 = 2          /* h*w*3 bytes: ImageSensor::saveImage tone curve with scale */
};
/* Copies the film out. scale multiplies the sums first (saveImage passes
 * brightness / spp, libSLR/Renderer/PTRenderer.cpp:84). SLRH_FILM_SRGB8 is
 * the pixel bytes ImageSensor::saveImage writes (libSLR/Core/ImageSensor.cpp:
 * 139-195) for a sensor holding the film's sums rounded to FP32, with its FP32
 * and double arithmetic, row 0 first, R, G, B. */
SLRH_API int slrh_film_read(const slrh_film* f, int format, double scale, void* dst, size_t dst_bytes);
/* Replaces the film's sums with w*h*16 doubles (pixel-major, 16 strata each):
 * to resume accumulating onto saved sums, or to encode sums made elsewhere. */
SLRH_API int slrh_film_load(slrh_film* f, const double* strata, size_t bytes);
/* The BMP file ImageSensor::saveImage(path, scale) writes for the film
 * (sensor sensitivity `sensitivity`, infinity counting as 1): saveBMP's
 * 54-byte header and bottom-up BGR rows of 3 * width + width % 4 bytes
 * (libSLR/Helper/bmp_exporter.cpp:15-54; the reference leaves the row padding
 * uninitialised, here it is zero). *size receives the file size; dst may be
 * NULL to query it. */
SLRH_API int slrh_film_bmp(const slrh_film* f, float scale, float sensitivity, void* dst, size_t dst_bytes,
                           size_t* size);

/* Renders into `film` (accumulating onto what it holds). Replaces
 * PTRenderer::render (libSLR/Renderer/PTRenderer.cpp:26-101). stats, if not
 * NULL, receives 27 doubles: [0] rays traced, [1] samples, [2] device seconds,
 * [3] non-finite samples dropped, [4] wavefront iterations, [5] traversal
 * kernel seconds (extension and shadow rays traced together), [6] traversal
 * launches, [7] extension rays, [8] shading kernel seconds, [9] seconds of
 * the other kernels (regeneration, compaction), [10] shadow rays, [11] QBVH
 * nodes visited and [12] triangles tested by extension rays, [13] nodes and
 * [14] triangles by shadow rays, [15] traversal variant used, [16] scattering
 * events in the medium and [17] light samples attenuated by it (SLRH_VOL_PT),
 * [18] seconds of the medium tracking kernel, and its [19] density lookups
 * (tentative collisions), [20] majorant cells entered, [21] rays tracked
 * and [22] tracking walks stopped by the step guard (2^16 steps; expected 0),
 * [23] QBVH nodes read from the traversal's LDS copy of the top of the tree
 * (included in [11] and [13]), [24] instances entered by the traversal,
 * [25] workgroup size of the surface shading kernels (chosen by the
 * autotuner with tune_trace 0), [26] slots (live paths) per pixel.
 * Kernel seconds are 0 unless flags bit 0. */
SLRH_API int slrh_render(slrh_scene* s, const slrh_render_settings* settings, slrh_film* film,
                         double* stats);

/* Progressive rendering: renders settings->spp samples per pixel into
 * `film` in passes of 1, 1, 2, 4, ... samples (the last pass cut to spp) and
 * calls snapshot(user, film, samples_done) after each pass, so the film
 * holds 1, 2, 4, 8, ... samples per pixel at the calls, where
 * PTRenderer::render saves its 000.bmp, 001.bmp, ... snapshots
 * (libSLR/Renderer/PTRenderer.cpp:77-94; the host saves one with
 * slrh_film_bmp(film, brightness / samples_done, ...)). A non-zero return
 * from snapshot stops the render there. Pass k draws from seed +
 * 0x9E3779B9 * k, so the final film is a valid render of spp samples, not
 * the same bits as one slrh_render call. stats summed over the passes
 * (stats[15], the traversal variant, averaged). snapshot may be NULL. */
typedef int (*slrh_snapshot_fn)(void* user, const slrh_film* film, uint32_t samples_done);
SLRH_API int slrh_render_progressive(slrh_scene* s, const slrh_render_settings* settings, slrh_film* film,
                                     slrh_snapshot_fn snapshot, void* user, double* stats);

/* One frame of a batch: camera (as slrh_camera_perspective), resolution,
 * samples per pixel, seed and the film it accumulates into. */
typedef struct slrh_frame {
    float l2w[16];
    float sensitivity, aspect, fov_y, lens_radius, img_dist, obj_dist;
    uint32_t width, height, spp, seed;
    slrh_film* film;
} slrh_frame;

/* Renders n frames of the scene in one wavefront: their paths share one pool
 * of path slots and every kernel launch (PTRenderer::render per frame,
 * libSLR/Renderer/PTRenderer.cpp:26-101, batched). settings supplies
 * max_depth, integrator, slots_per_pixel, flags and tune_trace; width,
 * height, spp and seed come from the frames. A frame's film holds the same
 * result as rendering it alone with the same seed and slots_per_pixel.
 * stats as for slrh_render, summed over the batch. */
SLRH_API int slrh_render_batch(slrh_scene* s, const slrh_frame* frames, uint32_t n,
                               const slrh_render_settings* settings, double* stats);

/* ------------------------------------------------------------ serving */

/* A render server over one resident scene (no counterpart in the reference,
 * which renders one frame per PTRenderer::render call). Requests queue in
 * order and are admitted while their path state fits the server's budget;
 * a scheduler thread keeps one wavefront running and lets requests join and
 * leave it between iteration blocks (continuous batching). A request's film
 * holds what slrh_render_batch gives for that frame with the same seed and
 * slots_per_pixel. The scene must outlive the server; films must outlive
 * their requests (until DONE, CANCELLED or FAILED). */
typedef struct slrh_server slrh_server;

typedef struct slrh_server_config {
    uint64_t path_state_bytes;   /* device memory for path state; bounds the live paths (~418 B each) */
    uint32_t tune_trace;         /* traversal kernel variant + 1 (0 = the one tuned by a render) */
    uint32_t reserved;           /* must be 0 */
} slrh_server_config;

typedef struct slrh_request {
    slrh_frame frame;            /* camera, resolution, spp, seed, film */
    uint32_t max_depth;          /* 0 = 100 */
    uint32_t slots_per_pixel;    /* 0 = chosen at admission from the load */
} slrh_request;

enum slrh_request_state {
    SLRH_REQ_QUEUED = 0, SLRH_REQ_RUNNING = 1, SLRH_REQ_DONE = 2, SLRH_REQ_CANCELLED = 3, SLRH_REQ_FAILED = 4
};

SLRH_API int slrh_server_create(slrh_scene* s, const slrh_server_config* cfg, slrh_server** out);
/* Stops the scheduler; requests not yet settled are abandoned. */
SLRH_API int slrh_server_destroy(slrh_server* sv);
SLRH_API int slrh_submit(slrh_server* sv, const slrh_request* req, uint64_t* ticket);
SLRH_API int slrh_poll(slrh_server* sv, uint64_t ticket, int* state);
/* Waits until the request settles or timeout_ms passes (< 0: no limit). */
SLRH_API int slrh_wait(slrh_server* sv, uint64_t ticket, int timeout_ms, int* state);
/* A cancelled request's film holds an unspecified part of its samples. */
SLRH_API int slrh_cancel(slrh_server* sv, uint64_t ticket);
/* stats[8]: completed requests, queued, running, p50 and p99 latency (ms,
 * submit to done), mean latency (ms), samples filed, scheduler busy seconds. */
SLRH_API int slrh_server_stats(slrh_server* sv, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* SLR_HIP_H */
