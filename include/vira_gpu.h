/* vira_gpu.h -- C-ABI of the MI355X (gfx950) ray-tracing hot path.
 *
 * This is the drop-in boundary below vira's renderer API.  In the reference
 * (crgnam/vira v0.8.3) the hot path is reached through
 *   CPUPathTracer::render(Camera&, Scene&)      include/vira/rendering/cpu_path_tracer.hpp:50
 * and, below it, the polymorphic acceleration interface
 *   BLAS::intersect / TLAS::intersect / TLAS::init
 *                                               include/vira/rendering/acceleration/blas.hpp:20-37
 *                                               include/vira/rendering/acceleration/tlas.hpp:19-31
 * whose working implementation is Embree (embree_blas.ipp:14-71,
 * embree_tlas.ipp:77-173).  Every entry point below names the reference
 * interface it replaces.  Plain pointers and sizes only; no exceptions cross
 * this boundary: each call returns a status and vira_gpu_last_error() gives
 * the message (thread-local).  The C++ wrapper (include/vira_gpu.hpp) turns
 * non-OK statuses into std::runtime_error, as the reference throws
 * (vira_tlas.ipp:33, cpu_path_tracer.ipp:121).
 *
 * Pointers: arguments named *_host are host memory; arguments named *_dev are
 * device memory (hipMalloc / torch CUDA tensors) on the context's device.
 * Every call is enqueued on the context's stream and is synchronous unless
 * stated otherwise.
 */
#ifndef VIRA_GPU_H
#define VIRA_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VIRA_GPU_NUM_BINS 8 /* SpectralData<8, UNIFORM_RANGE, 400, 1000> */
#define VIRA_GPU_MAX_LIGHTS 8

typedef enum {
    VIRA_GPU_OK = 0,
    VIRA_GPU_ERR_INVALID = 1,     /* bad argument (null pointer, size, id)      */
    VIRA_GPU_ERR_HIP = 2,         /* HIP runtime error (launch, fault, copy)    */
    VIRA_GPU_ERR_OOM = 3,         /* device allocation failed                   */
    VIRA_GPU_ERR_STATE = 4,       /* call order (e.g. render before BLAS build) */
    VIRA_GPU_ERR_UNSUPPORTED = 5  /* valid request this build does not handle   */
} vira_gpu_status;

typedef struct vira_gpu_ctx vira_gpu_ctx;

typedef struct {
    int device;   /* HIP device ordinal */
    void* stream; /* hipStream_t to enqueue on; NULL = a stream owned by the context */
} vira_gpu_opts;

/* DEM grid -> world projection.  kind 0: flat grid x = (i-cx)s, y = (cy-j)s,
 * z = h (the non-OGR branch of DEMProjection::transformToWorld,
 * dem_projection.ipp:130-141).  kind 1: gnomonic tile on a sphere of the
 * given radius centred on +x of the body frame.  Positions are formed in FP64
 * and `offset` (the rebasing origin, e.g. the camera position) is subtracted
 * before the FP32 store. */
typedef struct {
    int kind;
    double spacing;
    double radius;
    double center[3];
    double offset[3];
} vira_gpu_projection;

typedef struct {
    int max_leaf;   /* triangles per leaf, 1..15 (default 4)  */
    int bins;       /* SAH bins per node, 4..32 (default 16)  */
    int width;      /* 2 (binary, 64 B nodes) or 4 (collapsed, 128 B nodes; default) */
    float traversal_cost; /* SAH cost of a node visit relative to a triangle test (0 = 1) */
    int method;     /* 0 auto (grid for a mesh made from a DEM grid at width 4, else binned SAH);
                       1 binned SAH; 2 grid: for a mesh made from a DEM grid (vira_gpu_dem_mesh,
                       vira_gpu_qld_read) at width 4, the triangles in the Morton order of their grid
                       cells, the largest power of two up to max_leaf (and 8) to a leaf, under a
                       complete four-wide tree (each node a 2x2 block of its children on a square
                       power-of-two grid), boxes reduced bottom up */
    int compress;   /* width 4: also keep a copy of the nodes with 8-bit quantised child boxes (64 B per
                       node, decoded outward-rounded), which vira_gpu_render's per-pixel integrator and
                       vira_gpu_trace_rays traverse when every BLAS of the scene has one */
} vira_gpu_bvh_opts;

typedef struct {
    uint64_t num_nodes;      /* internal nodes of the binary tree      */
    uint64_t num_wide_nodes; /* nodes of the traversal layout (Node2 or Node4) */
    uint32_t width;          /* 2 or 4 */
    uint32_t wide_depth;     /* depth of the traversal layout */
    uint64_t num_leaves;
    uint32_t max_depth;
    uint32_t max_leaf_size;
    double sah_cost;         /* sum over internal nodes of area/root area, plus leaf tris*area/root area */
    double build_ms;         /* device time of the build, HIP events   */
    uint64_t bytes;          /* device bytes of nodes + leaf triangles */
    int32_t method;          /* the build that ran: 1 binned SAH, 2 grid */
    double wide_sah_cost;    /* width 4: sum over Node4s of area/root area, plus leaf tris*area/root area
                                (comparable between the two builds; sah_cost is the binary tree's, binned
                                SAH only) */
    uint64_t compressed_bytes; /* device bytes of the quantised node copy (0 without opts.compress) */
} vira_gpu_bvh_stats;

/* Light: kind 0 = PointLight (point_light.ipp:10-38), value = spectral
 * intensity P/(4 pi); kind 1 = SphereLight (sphere_light.ipp:16-125), value =
 * spectral radiance.  Position in the rebased frame (FP64). */
typedef struct {
    int kind;
    double position[3];
    float radius;
    float value[VIRA_GPU_NUM_BINS];
} vira_gpu_light;

/* Lambertian material (lambertian.ipp:9-15).  Albedo = vertex albedo x the
 * material's albedo map sampled bilinearly at the hit's uv (Image::sampleUVs,
 * image.ipp:589-670; Material::getAlbedo, material.ipp:114-116): either a
 * grey texture times albedo[k] (tex_channels 1) or a spectral albedo map of 8
 * bands per texel (tex_channels 8, Image<TSpectral>; albedo[] unused), or
 * albedo[k] alone without a texture.  An optional tangent-space normal map
 * (Material::getNormal, material.ipp:129-143) replaces the shading normal by
 * normalize([tangent bitangent N] * map(uv)). */
typedef struct {
    float albedo[VIRA_GPU_NUM_BINS];
    const float* texture_dev; /* row-major i + j*W, tex_channels floats per texel, or NULL */
    int tex_width;
    int tex_height;
    int smooth_shading;       /* Mesh::smoothShading (mesh.hpp:144) */
    int tex_channels;         /* 1 (or 0) grey, 8 spectral */
    const float* normal_map_dev; /* row-major, 3 floats (x, y, z in the tangent frame) per texel, or NULL */
    int nmap_width;
    int nmap_height;
    uint32_t material_id;     /* Material::getID, written to the material_id pass */
} vira_gpu_material;

/* Pinhole camera: kinv is GLM mat3x2 K^-1 stored [col][row] (camera.ipp:1526-1568),
 * rot row-major with columns = camera x, y, z axes (camera.ipp:1455-1500).
 * position is in the rebased frame; z_dir = +1 (or -1 for Blender frames). */
typedef struct {
    int width;
    int height;
    float kinv[6];
    float rot[9];
    double position[3];
    float z_dir;
    float view_normal[9]; /* Camera::getViewNormalMatrix (camera.hpp:190), row-major, for the
                             normal_camera pass; all zero = the transpose of rot */
} vira_gpu_camera;

/* CPUPathTracerOptions (cpu_path_tracer.hpp:30-43) plus the new controls. */
typedef struct {
    int samples;         /* spp */
    int bounces;
    int russian_roulette;
    int rr_start;        /* first bounce whose continuation is subject to RR */
    uint32_t seed;
    int integrator;      /* 0 auto (= 1), 1 fused per-pixel, 2 wavefront queues */
    int show_background; /* misses add throughput x background (cpu_path_tracer.ipp:257-270) */
    int adaptive_sampling;       /* cpu_path_tracer.ipp:198-227 (per-pixel integrator) */
    int samples_per_batch;       /* 0 = 30 */
    float sampling_tolerance;    /* 0 = 0.05 */
    int samples_to_detect_miss;  /* 0 = 30 */
    int count_visits;    /* per-pixel integrator: count BVH visits into the stats (a separate kernel;
                            one light or none, no ambient light, background, adaptive sampling or
                            material maps; the image is the same) */
} vira_gpu_render_opts;

/* Render passes (passes.hpp:51-95); every pointer optional, device memory,
 * pixel index i + j*W. */
typedef struct {
    float* radiance_dev;   /* W*H*8 total radiance            */
    float* direct_dev;     /* W*H*8 bounce-0 radiance         */
    float* indirect_dev;   /* W*H*8 total - direct            */
    float* albedo_dev;     /* W*H*8 mean first-hit albedo     */
    float* normal_dev;     /* W*H*3 normalised first-hit normal */
    float* depth_dev;      /* W*H mean first-hit distance (inf = no hit) */
    float* alpha_dev;      /* W*H hits / spp                  */
    uint32_t* tri_id_dev;  /* W*H first-hit triangle (0xffffffff = none) */
    uint32_t* instance_id_dev; /* W*H first-hit instance: index into the TLAS (0xffffffff = none) */
    float* normal_camera_dev;  /* W*H*3 normalize(view normal matrix x normal) (0 where no hit) */
    uint32_t* mesh_id_dev;     /* W*H first-hit mesh id (0xffffffff = none) */
    uint32_t* material_id_dev; /* W*H first-hit material id (vira_gpu_material.material_id) */
} vira_gpu_outputs;

/* Traversal work of one kind of ray, summed over the render (opt-in visit
 * counters; the reference's Vulkan path tracer keeps debug counters of the same
 * kind, ray_generation.glsl:274-276). */
typedef struct {
    uint64_t nodes;         /* internal BVH nodes fetched (each tests all its child boxes) */
    uint64_t leaves;        /* leaves entered */
    uint64_t triangles;     /* ray-triangle tests */
    uint64_t stack_spills;  /* pushes past the 8 stack entries kept in LDS (entry moved to private memory) */
    uint64_t tlas_nodes;    /* TLAS nodes whose box was tested (instanced scenes) */
    uint32_t max_stack;     /* deepest traversal stack of any ray */
    uint64_t wave_node_steps; /* iterations of the node loop summed over the waves (a wave steps until its
                                 last lane is done): SIMD use of the loop = nodes / (64 x this) */
} vira_gpu_visit_counts;

typedef struct {
    uint64_t primary_rays;
    uint64_t shadow_rays;
    uint64_t bounce_rays;
    double total_ms;         /* device time of the whole render (HIP events) */
    double trace_closest_ms; /* sum over closest-hit traversal launches */
    double trace_shadow_ms;  /* sum over any-hit traversal launches */
    double shade_ms;         /* sum over shading + accumulation launches */
    uint32_t trace_closest_launches;
    uint32_t trace_shadow_launches;
    int integrator;          /* 1 fused, 2 wavefront */
    double fused_ms;         /* device time of the fused renderer launch(es) */
    uint32_t fused_launches;
    uint64_t closest_hits;   /* camera and bounce rays that hit (shaded surface points; per-pixel renderer) */
    int visits_counted;      /* 1 when opts.count_visits filled the two fields below */
    vira_gpu_visit_counts closest_visits;  /* camera and bounce rays (closest hit) */
    vira_gpu_visit_counts shadow_visits;   /* shadow rays (any hit) */
} vira_gpu_render_stats;

/* Fixed-visibility direct illumination (the differentiable render).  New: the
 * reference has no differentiable renderer; the forward model is the bounce-0
 * direct-light term of processIntersection (cpu_path_tracer.ipp:284-428). */
typedef struct {
    int samples;            /* spp >= 1; sample 0 through the pixel corner */
    uint32_t seed;
    uint64_t record_budget; /* bytes of hit records (16 per pixel and sample) the forward pass may keep
                               for the adjoint; above it the adjoint re-traces them */
} vira_gpu_di_opts;

typedef struct {
    uint64_t primary_rays;
    uint64_t shadow_rays;
    double ms;              /* device time of the call's kernels (HIP events) */
    double trace_ms;        /* adjoint: the part that re-traced hit records (0 when they were kept) */
    int32_t used_records;   /* adjoint: 1 = read the forward pass's records, 0 = re-traced */
    uint64_t record_bytes;  /* forward: bytes of records kept (0 = over budget) */
} vira_gpu_di_stats;

/* Batched rendering (new): many render requests of uneven image size, spp and
 * mesh in one persistent launch over a shared queue of 8x8 pixel tiles. */
typedef struct {
    int32_t mesh_id;             /* rendered mesh: its BLAS and material */
    vira_gpu_camera camera;
    vira_gpu_render_opts opts;   /* integrator is ignored: the per-pixel renderer runs every request */
    int32_t num_lights;          /* 0 .. VIRA_GPU_MAX_LIGHTS */
    vira_gpu_light lights[VIRA_GPU_MAX_LIGHTS];
    vira_gpu_outputs out;        /* caller-owned device passes, as for vira_gpu_render */
} vira_gpu_request;

typedef struct {
    int32_t bvh_width;        /* 2 or 4; every request's mesh must have a BLAS of this width */
    int32_t waves_per_block;  /* 1, 2 or 4 */
    int32_t waves_per_cu;     /* persistent waves per compute unit: the depth of the shared queue's consumers */
    int32_t tiles_per_grab;   /* tiles taken from the shared queue per atomic, 1..16 */
    int32_t order;            /* 0 = requests with the most rays per tile first, 1 = submission order */
} vira_gpu_batch_config;

typedef struct {
    uint64_t primary_rays;
    uint64_t shadow_rays;
    uint64_t bounce_rays;
    uint64_t tiles;
    double ms;                /* device time of the launch (HIP events) */
    int32_t graph_reused;     /* 1 = replayed the captured hipGraph of an identical batch layout */
    uint64_t closest_hits;    /* camera and bounce rays that hit (shaded surface points) */
} vira_gpu_batch_stats;

/* Sensor radiometry of Camera::calculateReceivedPower (camera.ipp:1175-1178). */
typedef struct {
    float optical_efficiency[VIRA_GPU_NUM_BINS]; /* eta per band (default 1) */
    float aperture_area;                         /* pi (f/N / 2)^2, m^2 (aperture.ipp:14-16, camera.ipp:82) */
} vira_gpu_sensor;

/* Point-spread function (PointSpreadFunction::makeKernel, psf.ipp): kind 0 =
 * GaussianPSF (gaussian_psf.ipp; sigmas in pixels per band, rotation angle in
 * degrees), kind 1 = AiryDiskPSF (airy_disk.ipp; focal length, aperture
 * diameter and pixel pitch in metres, band centre wavelengths in metres). */
typedef struct {
    int kind;
    float sigma_x[VIRA_GPU_NUM_BINS];
    float sigma_y[VIRA_GPU_NUM_BINS];
    float angle_deg;
    float focal_length;
    float aperture_diameter;
    float pixel_size[2];
    float wavelength[VIRA_GPU_NUM_BINS];
    int supersample;  /* sub-samples per kernel-pixel side (initKernels uses 20) */
} vira_gpu_psf;

/* Sensor tail (Camera::simulateSensor / getPhotonCounts, camera.ipp:1221-1340;
 * Photosite::exposePixel, photosite.ipp; NoiseModel::simulate, noise_model.ipp). */
typedef struct {
    float exposure_time;                          /* s */
    float photon_energy[VIRA_GPU_NUM_BINS];       /* J per photon at each band centre */
    float quantum_efficiency[VIRA_GPU_NUM_BINS];
    float gain;                                   /* ADU per electron */
    float max_adu;                                /* linear_scale_factor * (2^bit_depth - 1) */
    int32_t photon_noise;                         /* Poisson photon noise when >= 1 photon in the pixel */
    int32_t noise_model;                          /* dark current, readout noise, fixed pattern */
    float dark_current;                           /* electrons / s (Poisson rate) */
    float readout_noise_mean, readout_noise_std;
    float horizontal_scale, vertical_scale;
    int32_t horizontal_pattern_period, vertical_pattern_period;
    uint32_t seed;                                /* counter-based RNG key */
    const float* filter_mosaic_dev;  /* W*H*8 colour filter array applied to the received energy
                                        (Camera::hasFilterMosaic, camera.ipp:1293-1306), or NULL */
} vira_gpu_sensor_sim;

/* Point sources for the unresolved renderer (CPUUnresolvedRenderer::render,
 * cpu_unresolved_renderer.ipp:25-200): PSF kernels of ascending odd sizes
 * (PointSpreadFunction::initKernels, psf.ipp), the edge maxima that choose
 * between them (getKernel), the minimum power below which a source is
 * skipped, and the camera's intrinsic matrix K as GLM mat3x2 [col][row]
 * (camera.ipp:1555-1557) for Camera::projectCameraPoint (camera.ipp:1062-1081). */
typedef struct {
    int32_t count;                                 /* 0 (no PSF) .. 4 kernels */
    int32_t size[4];
    const float* kernel_dev[4];                    /* size*size*8 each, as vira_gpu_psf_kernel writes */
    float edge_max[4][VIRA_GPU_NUM_BINS];
    float minimum_power;                           /* W */
    float intrinsic[6];                            /* K: fx kyx | kxy fy | px py */
    int32_t cull_behind;                           /* 0: as the reference, 1: skip camera-frame z <= 0 */
} vira_gpu_star_psf;

/* EATWTOptions (cpu_denoise.hpp:9-18); TILE_SIZE has no counterpart. */
typedef struct {
    int32_t max_levels_direct;     /* 5 */
    int32_t max_levels_indirect;   /* 5 */
    float depth_threshold;         /* 0.2 */
    float normal_threshold;        /* 0.5 */
    float min_normal_weight;       /* 0.3 */
    float epsilon;                 /* 1e-6 */
} vira_gpu_denoise_opts;

/* --- context ------------------------------------------------------------ */
vira_gpu_status vira_gpu_create(vira_gpu_ctx** ctx, const vira_gpu_opts* opts);
vira_gpu_status vira_gpu_destroy(vira_gpu_ctx* ctx);
vira_gpu_status vira_gpu_set_stream(vira_gpu_ctx* ctx, void* stream);
vira_gpu_status vira_gpu_synchronize(vira_gpu_ctx* ctx);
const char* vira_gpu_last_error(void);
const char* vira_gpu_version(void);

/* --- synthetic inputs (value-noise fBm DEM, per-vertex albedo) ------------ */
vira_gpu_status vira_gpu_synth_fbm(vira_gpu_ctx* ctx, int width, int height, uint32_t seed,
                                   double amplitude, int octaves, double base_cells, float* heights_dev);
vira_gpu_status vira_gpu_synth_vertex_albedo(vira_gpu_ctx* ctx, size_t num_vertices, uint32_t seed,
                                             float* albedo_dev);

/* --- meshes ---------------------------------------------------------------- */
/* DEM::makeMesh (dem.ipp:138-211) + imageToIndexBuffer (image_utils.ipp:364-412)
 * + Mesh::calculateNormals (mesh.ipp:199-233), on the GPU.  heights_dev is
 * W*H row-major (i + j*W); grey_albedo_dev is W*H or NULL (albedo 1).
 * UVs are (i/(W-1), j/(H-1)). */
vira_gpu_status vira_gpu_dem_mesh(vira_gpu_ctx* ctx, const float* heights_dev, int width, int height,
                                  const vira_gpu_projection* proj, const float* grey_albedo_dev,
                                  int32_t* mesh_id);
/* Mesh(vertexBuffer, indexBuffer) (mesh.ipp:25-83) from host buffers.
 * albedo_host: nv*8 or NULL (1); uv_host: nv*2 or NULL (0); normals_host:
 * nv*3 or NULL (computed as Mesh::calculateNormals). */
vira_gpu_status vira_gpu_mesh_upload(vira_gpu_ctx* ctx, const float* positions_host, size_t num_vertices,
                                     const uint32_t* indices_host, size_t num_triangles,
                                     const float* albedo_host, const float* uv_host,
                                     const float* normals_host, int32_t* mesh_id);
vira_gpu_status vira_gpu_mesh_info(vira_gpu_ctx* ctx, int32_t mesh_id, size_t* num_vertices,
                                   size_t* num_triangles);
/* Copies the mesh back (positions nv*3, indices nt*3, normals nv*3); any may be NULL. */
vira_gpu_status vira_gpu_mesh_download(vira_gpu_ctx* ctx, int32_t mesh_id, float* positions_host,
                                       uint32_t* indices_host, float* normals_host);
vira_gpu_status vira_gpu_mesh_release(vira_gpu_ctx* ctx, int32_t mesh_id);
/* New heights for a mesh made by vira_gpu_dem_mesh (DEM::makeMesh,
 * dem.ipp:165-211, re-run on the same grid): positions and normals are
 * recomputed in place and a built BLAS is rebuilt with its last options.  The
 * set of non-finite heights (holes) must not change. */
vira_gpu_status vira_gpu_dem_update(vira_gpu_ctx* ctx, int32_t mesh_id, const float* heights_dev);
/* vira_gpu_dem_update with the BLAS refitted (vira_gpu_blas_refit) instead of
 * rebuilt: for the small height changes of an optimisation loop. */
vira_gpu_status vira_gpu_dem_update_refit(vira_gpu_ctx* ctx, int32_t mesh_id, const float* heights_dev);

/* --- acceleration ---------------------------------------------------------- */
/* Mesh::buildBVH -> EmbreeBLAS (mesh.ipp:396-418, embree_blas.ipp:14-71):
 * binned-SAH build on the GPU, or the grid build for DEM meshes (opts.method).
 * opts may be NULL. */
vira_gpu_status vira_gpu_blas_build(vira_gpu_ctx* ctx, int32_t mesh_id, const vira_gpu_bvh_opts* opts);
vira_gpu_status vira_gpu_blas_stats(vira_gpu_ctx* ctx, int32_t mesh_id, vira_gpu_bvh_stats* stats);
/* New (the reference rebuilds, embree_blas.ipp:14-71): a built BLAS refitted
 * to its mesh's current vertex positions, same triangles.  The tree is kept;
 * the leaf triangles and every box are recomputed bottom up, so hits stay
 * exact while traversal cost drifts with the deformation.  A 2-wide BLAS is
 * rebuilt with its last options.  stats.build_ms becomes the refit's time;
 * the other statistics describe the tree as built. */
vira_gpu_status vira_gpu_blas_refit(vira_gpu_ctx* ctx, int32_t mesh_id);

typedef struct {
    int32_t mesh_id;
    double transform[16]; /* column-major 4x4 model matrix, FP64 */
} vira_gpu_instance;

/* Scene::buildTLAS -> EmbreeTLAS::newInstance/build (scene.ipp:712-812,
 * embree_tlas.ipp:77-115).  1..16 instances, each a built BLAS (all of one
 * BVH width) placed by an FP64 column-major model matrix (invertible affine).
 * One instance with the identity matrix renders in the caller's frame as
 * given.  Any other TLAS is rendered camera-relative: at each render a kernel
 * subtracts the FP64 camera position from every model matrix and light
 * position and rounds the results (and the inverses and normal matrices) to
 * FP32, so positions may be planetary- or AU-scale.  Instanced scenes use the
 * per-pixel renderer (at most one light); the direct-illumination adjoint
 * takes a single-instance TLAS. */
vira_gpu_status vira_gpu_tlas_build(vira_gpu_ctx* ctx, const vira_gpu_instance* instances_host, size_t n);

/* --- shading state -------------------------------------------------------- */
/* Scene::light_cache_ (scene.hpp:263-285). */
vira_gpu_status vira_gpu_lights_set(vira_gpu_ctx* ctx, const vira_gpu_light* lights_host, int n);
/* Mesh::material_cache_ (mesh.hpp:93). */
vira_gpu_status vira_gpu_material_set(vira_gpu_ctx* ctx, int32_t mesh_id, const vira_gpu_material* mat);
/* Scene::setAmbient / setBackgroundEmission (scene.hpp:90-113): ambient
 * spectral radiance (8 values, host; NULL or all zero = no ambient light),
 * added at the first hit as throughput x ambient x albedo
 * (cpu_path_tracer.ipp:389-396), and the background image (W*H*8 on the
 * device, kept by reference; NULL = zero) seen by rays that miss when
 * show_background is set.  Both are used by the per-pixel integrator. */
vira_gpu_status vira_gpu_environment_set(vira_gpu_ctx* ctx, const float* ambient_host, const float* background_dev,
                                         int bg_width, int bg_height);

/* --- rendering --------------------------------------------------------- */
/* CPUPathTracer::render (cpu_path_tracer.ipp:29-161) over the current TLAS. */
vira_gpu_status vira_gpu_render(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_render_opts* opts,
                                const vira_gpu_outputs* out, vira_gpu_render_stats* stats);

/* No counterpart in the reference: vira_gpu_render for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It renders with the per-pixel (fused)
 * integrator, as vira_gpu_render does by default, and writes the same passes
 * bit for bit; its scratch (tile accumulators, ray counters, light and
 * instance tables) is one fixed array per GPU that the library holds for the
 * life of the process (about 49 MiB), so no hipMalloc, hipFree or other
 * allocating call is made.  Limits: at most 4096 8x8 pixel tiles (e.g. 512 x
 * 512 pixels) and 64 instances (VIRA_GPU_ERR_INVALID beyond either); the
 * wavefront integrator returns VIRA_GPU_ERR_UNSUPPORTED.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (two contexts, two
 * threads) race. */
vira_gpu_status vira_gpu_render_noalloc(vira_gpu_ctx* ctx, const vira_gpu_camera* cam,
                                        const vira_gpu_render_opts* opts, const vira_gpu_outputs* out,
                                        vira_gpu_render_stats* stats);

/* Parity hook: Scene::intersect -> TLAS::intersect (scene.hpp:230-233,
 * embree_tlas.ipp:128-173) for a batch of rays against one mesh's BLAS.
 * rays_dev: n x 8 floats (ox oy oz 0 dx dy dz tmax); any_hit != 0 -> occlusion
 * within [0, tmax] (t = 0 where occluded, +inf otherwise).  Outputs n floats t
 * (+inf = miss), n uint32 triangle ids (0xffffffff = miss), n floats b1, b2
 * (Embree barycentrics u, v); any output may be NULL.  kernel_ms (may be NULL)
 * receives the traversal kernel's device time. */
vira_gpu_status vira_gpu_trace_rays(vira_gpu_ctx* ctx, int32_t mesh_id, const float* rays_dev, size_t n,
                                    int any_hit, float* t_dev, uint32_t* tri_id_dev, float* b1_dev,
                                    float* b2_dev, double* kernel_ms);

/* --- differentiable direct illumination (new) ----------------------------- */
/* Forward: radiance (W*H*8, as vira_gpu_outputs.radiance_dev) of the scene
 * mesh lit by point lights, one bounce-0 direct-light evaluation per sample
 * (identical to vira_gpu_render with bounces = 0).  Keeps the hit records
 * (triangle, barycentrics, per-light visibility) for vira_gpu_adjoint when
 * they fit in opts->record_budget. */
vira_gpu_status vira_gpu_di_render(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_di_opts* opts,
                                   float* radiance_dev, vira_gpu_di_stats* stats);
/* Adjoint of vira_gpu_di_render with visibility held fixed: from dL/dimage
 * (W*H*8) to dL/dheights (one per grid node of the scene mesh, which must come
 * from vira_gpu_dem_mesh; 0 at holes) and dL/dtexels (tex_width*tex_height of
 * the scene material); either output may be NULL and both are overwritten.
 * Uses the records of the preceding vira_gpu_di_render with the same camera,
 * samples and seed and an unchanged scene; re-traces them otherwise.  Point
 * lights only (VIRA_GPU_ERR_UNSUPPORTED for sphere lights). */
vira_gpu_status vira_gpu_adjoint(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_di_opts* opts,
                                 const float* dL_dimage_dev, float* dL_dheights_dev, float* dL_dtexture_dev,
                                 vira_gpu_di_stats* stats);

/* Received power P = eta * A * Omega_ij * L per pixel and band
 * (cpu_path_tracer.ipp:137-146 -> camera.ipp:1175-1178), Omega_ij the pixel's
 * solid angle (camera.ipp:1594-1666, FP64 on the device, stored FP32).
 * radiance_dev / power_dev: W*H*8; solid_angle_dev: W*H; power or solid angle
 * output may be NULL.  The solid angles are computed once per camera geometry
 * (size, K^-1, z sign) and kept on the context, as the reference computes them
 * when the camera is initialised. */
vira_gpu_status vira_gpu_received_power(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_sensor* sensor,
                                        const float* radiance_dev, float* power_dev, float* solid_angle_dev);

/* PointSpreadFunction::makeKernel (psf.ipp): size x size x 8 kernel, value of
 * kernel pixel (x, y) band k at kernel_dev[(x + y*size)*8 + k]; size odd >= 3. */
vira_gpu_status vira_gpu_psf_kernel(vira_gpu_ctx* ctx, const vira_gpu_psf* psf, int size, float* kernel_dev);
/* Image::convolve (image.ipp:993-1065, 1267-1330): per band, zero-padded,
 * out(x, y) = sum_{kx, ky} in(x + kx - K/2, y + ky - K/2) * kernel(kx, ky).
 * image_dev / out_dev: W*H*8 (out must not alias image).  Kernels below 256
 * taps are summed directly (the reference's spatial branch, same order);
 * larger ones go through 2-D FFTs (its FFT branch) when W and H plus K/2 fit
 * 8192, with the three buffers 8-byte aligned, and are summed directly for
 * larger images.  The kernel's spectrum is kept on the context for the next
 * call with the same kernel values and padded sizes. */
vira_gpu_status vira_gpu_convolve(vira_gpu_ctx* ctx, const float* image_dev, int width, int height,
                                  const float* kernel_dev, int ksize, float* out_dev);
/* Camera::simulateSensor: received power (W*H*8, watts) -> photon counts
 * (W*H*8, optional) and the normalised digital image (W*H, 0..1). */
vira_gpu_status vira_gpu_simulate_sensor(vira_gpu_ctx* ctx, const vira_gpu_sensor_sim* cfg, const float* power_dev,
                                         int width, int height, float* photons_dev, float* image_dev);

/* CPUUnresolvedRenderer (cpu_unresolved_renderer.ipp:25-200): n point sources
 * with global unit directions (n*3) and spectral irradiance (n*8, W/m^2 per
 * band) are turned to the camera frame (R^T d), projected with K, given the
 * received power eta * A * E (camera.ipp:1187-1191) and splatted with the
 * bilinear-sampled PSF response (Image::addImage, image.ipp:1074-1150) into
 * power_dev (W*H*8, accumulated: clear it first).  With no PSF (count 0)
 * a source's power replaces the value of pixel ((int) x, (int) y), the
 * last source in order winning (processRegion, :191-196).  Each pixel receives its
 * sources' responses in source order, as the reference's sequential loop
 * adds them, so the result is deterministic.  The reference projects every
 * source, so one behind the camera lands where its antipode would
 * (cull_behind = 0); cull_behind = 1 skips camera-frame z <= 0. */
vira_gpu_status vira_gpu_render_stars(vira_gpu_ctx* ctx, const vira_gpu_camera* cam, const vira_gpu_sensor* sensor,
                                      const vira_gpu_star_psf* psf, const float* directions_dev,
                                      const float* irradiance_dev, size_t n, float* power_dev);

/* Star (star.hpp:10-40) as a star quipu stores it (StarQuipu::write,
 * star_quipu.ipp:31-41: the raw 64-byte record): the initialised flag and
 * padding, catalogue right ascension and declination (radians), proper
 * motions (radians per year), V magnitude, temperature (K) and angular size
 * (sr). */
typedef struct {
    uint8_t initialized;
    uint8_t pad[7];
    double ra, de, pm_ra, pm_de, vmag, temperature, omega;
} vira_gpu_star;

/* StarCatalogue::makeStarLight (star_catalogue.ipp:55-78) for n stars in
 * device memory: the unit vector at ephemeris time et, seconds
 * (Star::getUnitVector, star.ipp:19-32, FP32 from FP64 angles), and the
 * irradiance omega * blackBodyRadiance(T, 100) per band (Star::getIrradiance,
 * star.ipp:13-16; spectral_data.ipp:613-629: a 100-point trapezoid of Planck's
 * law in FP64 over each band).  As in the reference, pm_ra (rad/yr) moves the
 * declination and pm_de the right ascension.  stars_dev 8-byte aligned;
 * directions_dev n*3, irradiance_dev n*8: the inputs of vira_gpu_render_stars. */
vira_gpu_status vira_gpu_star_light(vira_gpu_ctx* ctx, const vira_gpu_star* stars_dev, size_t n, double et,
                                    float* directions_dev, float* irradiance_dev);
/* StarQuipu::read (star_quipu.ipp:20-29): the stars of a .qsc file into host
 * memory.  With stars_host NULL only *count is set; otherwise up to capacity
 * records are copied and *count is the number in the file.  Host only: no
 * context or GPU is touched. */
vira_gpu_status vira_gpu_qsc_read(const char* path, vira_gpu_star* stars_host, size_t capacity, size_t* count);

/* denoiseSpectralRadianceEATWT (cpu_denoise.ipp:10-187), as
 * CPUPathTracer::render applies it when options.denoise is set
 * (cpu_path_tracer.ipp:124-134): direct_dev and indirect_dev (W*H*8) are
 * denoised in place, edge-stopped by depth_dev (W*H) and normal_dev (W*H*3,
 * the global first-hit normal), demodulated by albedo_dev (W*H*8); total_dev
 * (optional) receives direct + indirect.  The first pass of each level runs
 * along the image's j axis, the second along i (the reference's order on
 * square images).  Spectral images 16-byte aligned. */
vira_gpu_status vira_gpu_denoise(vira_gpu_ctx* ctx, const vira_gpu_denoise_opts* opts, int width, int height,
                                 float* direct_dev, float* indirect_dev, const float* albedo_dev,
                                 const float* depth_dev, const float* normal_dev, float* total_dev);

/* --- quipu DEM pyramids (.qld) -------------------------------------------- */
/* DEMQuipu (dem_quipu.hpp:34-81, dem_quipu.ipp:29-236): a file of DEM levels
 * of decreasing resolution, each with its GSD.  The reader parses the file
 * and decodes its LZ4 blocks on the host and builds the level's mesh on the
 * GPU, ready for vira_gpu_blas_build. */
typedef struct vira_gpu_qld vira_gpu_qld;
typedef struct {
    int32_t levels;
    int32_t has_albedos;
    double transformation[16];   /* model matrix, column-major (readTransformation) */
    float normal[3];             /* normal cone of the coarsest level */
    float cone_angle;
} vira_gpu_qld_info;
typedef struct {
    int32_t read_albedos;        /* DEMQuipuReaderOptions::readAlbedos */
    float default_albedo;        /* used when the file has no albedo or read_albedos = 0 (0.06) */
    float albedo_profile[VIRA_GPU_NUM_BINS];  /* grey albedo images are multiplied by this (1) */
    double offset[3];            /* subtracted from OGR (map-projected) positions, as makeVertexBuffer does */
} vira_gpu_qld_read_opts;
typedef struct {
    int32_t level;
    double gsd;
    int32_t width, height;
    int32_t height_bits;         /* 8, 16 (fixed point) or 32 (float) */
    int32_t compressed;          /* heights stored as an LZ4 block */
    int32_t albedo_type;         /* -1 none, 0 constant, 1 grey image, 2 eight channels */
    int32_t is_ogr;
    double read_ms, decode_ms, build_ms;  /* file read + parse, of which LZ4 decode; device work */
} vira_gpu_qld_level;
/* The LZ4 block decoder the reader uses (LZ4_decompress_safe's contract,
 * quipu_io.ipp:219-235): returns the decoded size, or -1 for a malformed or
 * truncated block or one that does not fit in dst_capacity.  Host memory. */
int64_t vira_gpu_lz4_decode_block(const void* src, size_t src_size, void* dst, size_t dst_capacity);
/* DEMQuipu::DEMQuipu (dem_quipu.ipp:29-56): header and level table. */
vira_gpu_status vira_gpu_qld_open(const char* path, vira_gpu_qld** qld, vira_gpu_qld_info* info);
vira_gpu_status vira_gpu_qld_close(vira_gpu_qld* qld);
vira_gpu_status vira_gpu_qld_level_gsd(const vira_gpu_qld* qld, int32_t level, double* gsd);
/* readBuffers' choice (dem_quipu.ipp:137-151): the coarsest level whose GSD
 * is <= required_gsd, else level 0. */
vira_gpu_status vira_gpu_qld_select(const vira_gpu_qld* qld, double required_gsd, int32_t* level);
/* DEMQuipu::readBuffers (dem_quipu.ipp:131-236) for one level: the mesh
 * (DEM::makeVertexBuffer / makeIndexBuffer, dem.ipp:138-211, with the level's
 * DEMProjection: planar, or OGR equirectangular, or OGR polar stereographic on
 * a sphere) in a new mesh id; heights_dev
 * (W*H, optional) and albedo_dev (W*H*8, optional) receive the decoded
 * samples; mesh_id may be NULL to decode only.  Other OGR projections, and
 * polar stereographic on an ellipsoid, return VIRA_GPU_ERR_UNSUPPORTED. */
vira_gpu_status vira_gpu_qld_read(vira_gpu_ctx* ctx, vira_gpu_qld* qld, int32_t level,
                                  const vira_gpu_qld_read_opts* opts, int32_t* mesh_id, float* heights_dev,
                                  float* albedo_dev, vira_gpu_qld_level* info);

/* --- batched rendering (new) --------------------------------------------- */
/* Renders n requests (CPUPathTracer::render semantics per request,
 * cpu_path_tracer.ipp:29-161) in one persistent, hipGraph-captured launch;
 * each request's passes equal what vira_gpu_render writes for it (same
 * integrator, same sample order).  cfg NULL = the configuration chosen by
 * the last vira_gpu_batch_autotune, else a default. */
vira_gpu_status vira_gpu_batch_submit(vira_gpu_ctx* ctx, const vira_gpu_request* requests_host, size_t n,
                                      const vira_gpu_batch_config* cfg, vira_gpu_batch_stats* stats);
/* Times candidate configurations (BVH width, waves per block, waves per CU,
 * tiles per grab) on every sample_stride-th tile of each request (best of
 * two runs), re-times the three fastest on all tiles when sample_stride > 1,
 * keeps the fastest of those for later submits and returns it (rebuilding
 * the requests' BLASes at its BVH width). */
vira_gpu_status vira_gpu_batch_autotune(vira_gpu_ctx* ctx, const vira_gpu_request* requests_host, size_t n,
                                        int sample_stride, vira_gpu_batch_config* best, double* best_ms);

#ifdef __cplusplus
}
#endif

#endif /* VIRA_GPU_H */
