/*
 * rmh.h -- C ABI of librmh.so, the MI355X (gfx950) ray-casting simulation library.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an int
 * status (RMH_OK == 0).  On failure the thread-local message is available from
 * rmh_last_error().  Device pointers ("_dev") are HIP device allocations; host
 * pointers are ordinary memory.  Streams are hipStream_t passed as void* (NULL means the
 * legacy default stream).
 *
 * The entry points replace the following pieces of Rmagine (uos/rmagine v2.4.2):
 *   rmh_map_build        OptixScene::commit GAS build      src/rmagine_optix/src/map/optix/OptixScene.cpp:230-379
 *                        EmbreeScene::commit               src/rmagine_embree/src/map/embree/EmbreeScene.cpp:187-191
 *                        face normals                      src/rmagine_cuda/src/map/mesh_preprocessing.cu:6-43
 *   rmh_sensor_create    Sphere/Pinhole/O1Dn SimulatorOptix::setModel (uploads the model)
 *                                                          src/rmagine_optix/src/simulation/SphereSimulatorOptix.cpp:20-60
 *                                                          src/rmagine_optix/src/simulation/O1DnSimulatorOptix.cpp:41-83
 *   rmh_simulate         *SimulatorOptix::simulate(Memory<Transform,VRAM_CUDA>, Bundle&)
 *                                                          src/rmagine_optix/include/rmagine/simulation/SphereSimulatorOptix.tcc:41-100
 *                        with the Embree result contract   src/rmagine_embree/include/rmagine/simulation/SphereSimulatorEmbree.tcc:50-254
 *   rmh_malloc/...       VRAM_CUDA memory tag              src/rmagine_cuda/src/types/MemoryCuda.cu:13-89
 */
#ifndef RMH_H
#define RMH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMH_VERSION_MAJOR 0
#define RMH_VERSION_MINOR 1

/* status codes */
#define RMH_OK              0
#define RMH_ERR_INVALID     1  /* bad argument (null pointer, size, kind)         */
#define RMH_ERR_HIP         2  /* a HIP runtime call failed                         */
#define RMH_ERR_NOMAP       3  /* simulate without a map (reference: runtime_error) */
#define RMH_ERR_ALLOC       4  /* device allocation failed                          */
#define RMH_ERR_UNSUPPORTED 5  /* feature or size not supported                     */

/* ---- POD layouts, byte-identical to Rmagine's types ------------------------------ */

/* rmagine::Vector3 (src/rmagine_core/include/rmagine/math/types/Vector3.hpp:54-56) */
typedef struct { float x, y, z; } rmh_vec3;
/* rmagine::Quaternion (Quaternion.hpp:61-64) */
typedef struct { float x, y, z, w; } rmh_quat;
/* rmagine::Transform: 32 B {q, t, stamp} (Transform.hpp:63-68) */
typedef struct { rmh_quat R; rmh_vec3 t; uint32_t stamp; } rmh_transform;
/* rmagine::Interval (src/rmagine_core/include/rmagine/types/sensor_models.h:55-70) */
typedef struct { float min, max; } rmh_interval;
/* rmagine::DiscreteInterval (sensor_models.h:72-110) */
typedef struct { float min, inc; uint32_t size; } rmh_discrete_interval;
/* rmagine::SphericalModel, 32 B (sensor_models.h:122-214) */
typedef struct { rmh_discrete_interval phi, theta; rmh_interval range; } rmh_spherical_model;
/* rmagine::PinholeModel, 32 B (sensor_models.h:218-316) */
typedef struct { uint32_t width, height; rmh_interval range; float f[2]; float c[2]; } rmh_pinhole_model;
/* rmagine::O1DnModel_ flattened (sensor_models.h:338-418): dirs is a device array of
 * width*height directions, copied by rmh_sensor_create. */
typedef struct {
  uint32_t width, height; rmh_interval range; rmh_vec3 orig; uint32_t _pad;
  const rmh_vec3* dirs_dev;
} rmh_o1dn_model;
/* rmagine::OnDnModel_ flattened (sensor_models.h:422-503): origs and dirs are device
 * arrays of width*height entries, copied by rmh_sensor_create. */
typedef struct {
  uint32_t width, height; rmh_interval range;
  const rmh_vec3* origs_dev; const rmh_vec3* dirs_dev;
} rmh_ondn_model;

enum { RMH_MODEL_SPHERICAL = 0, RMH_MODEL_PINHOLE = 1, RMH_MODEL_O1DN = 2, RMH_MODEL_ONDN = 3 };

/* Result buffers of one simulate call (src/rmagine_core/include/rmagine/simulation/
 * SimulationResults.hpp:55-137).  Each is a device array of n_poses*width*height
 * entries indexed glob_id = pose*W*H + row*W + col, or NULL when the attribute is not
 * requested (the reference enables an attribute iff its buffer size > 0,
 * src/rmagine_embree/include/rmagine/simulation/embree_common.h:79-139). */
typedef struct {
  uint8_t*  hits;
  float*    ranges;
  rmh_vec3* points;
  rmh_vec3* normals;
  uint32_t* face_ids;
  uint32_t* geom_ids;
  uint32_t* object_ids;
} rmh_outputs;

typedef struct rmh_map rmh_map;
typedef struct rmh_sensor rmh_sensor;

/* BVH build parameters (NULL = defaults). */
typedef struct {
  uint32_t max_leaf_size;   /* 1..8, default 4                          */
  float    sah_cost_node;   /* SAH traversal cost, default 1.0          */
  float    sah_cost_tri;    /* SAH intersection cost, default 1.5       */
  uint32_t flags;           /* RMH_BVH_* bits, default 0                */
} rmh_bvh_params;
/* Also collapse the binary tree into a 4-wide tree with 8-bit quantised child boxes
 * (64 B nodes) and let the ray kernels (simulate, batch, pose gradient) traverse that one.
 * Results are identical; on the measured workloads the binary tree is faster, so it is the
 * default.  Closest-point queries and refit always use the binary tree. */
#define RMH_BVH_WIDE 1u
/* Skip the SAH refinement of the radix tree (three rounds of treelet restructuring,
 * Karras and Aila, HPG 2013): faster build, slower traversal. */
#define RMH_BVH_NO_REFINE 2u

typedef struct {
  uint32_t n_vertices, n_faces, n_meshes;
  uint32_t n_nodes;         /* reachable interior nodes                   */
  uint32_t n_leaves;        /* reachable leaves                           */
  uint32_t max_depth;       /* deepest leaf (root = 0)                    */
  uint64_t device_bytes;    /* nodes + triangles + normals                */
  float    build_ms;        /* device time of the build                   */
  float    sah_cost;        /* SAH cost of the tree (root-area relative)  */
  uint32_t n_wide_nodes;    /* nodes of the 4-wide tree (0: binary only)   */
  uint32_t wide_levels;     /* most 4-wide nodes on a root-to-leaf path    */
} rmh_map_info;

/* ---- runtime ---------------------------------------------------------------------- */
const char* rmh_last_error(void);
int rmh_version(void);
int rmh_get_device_count(int* count);
int rmh_set_device(int device);
int rmh_malloc(void** ptr_dev, size_t bytes);
int rmh_free(void* ptr_dev);
/* kind: 0 host->device, 1 device->host, 2 device->device, 3 host->host.  Blocking, like
 * the reference's cudaMemcpy copies (MemoryCuda.cu:25-29). */
int rmh_memcpy(void* dst, const void* src, size_t bytes, int kind);
int rmh_memset(void* ptr_dev, int value, size_t bytes);
int rmh_stream_create(void** stream);
int rmh_stream_destroy(void* stream);
int rmh_stream_synchronize(void* stream);
int rmh_device_synchronize(void);

/* ---- map (mesh + BVH in HBM) ------------------------------------------------------- */
/* vertices_dev: n_vertices xyz float triples; faces_dev: n_faces uint32 index triples
 * (indices into the concatenated vertex array).  mesh_face_offsets (host, n_meshes+1
 * entries, or NULL for a single mesh) splits the faces into geometries: face f with
 * offsets[g] <= f < offsets[g+1] reports geom_id = object_id = g and face_id =
 * f - offsets[g], which is what Embree reports for meshes attached to one scene
 * (EmbreeScene.cpp:54-71).  The map copies the mesh; the inputs may be freed after the
 * call returns.  The build runs on `stream` and is synchronised before returning.
 * At most 2^26 faces (the ray kernels address BVH nodes with 32-bit byte offsets):
 * larger maps are refused with RMH_ERR_UNSUPPORTED before any device work. */
int rmh_map_build(const float* vertices_dev, uint32_t n_vertices,
                  const uint32_t* faces_dev, uint32_t n_faces,
                  const uint32_t* mesh_face_offsets, uint32_t n_meshes,
                  const rmh_bvh_params* params, void* stream, rmh_map** out);
int rmh_map_destroy(rmh_map* map);
int rmh_map_get_info(const rmh_map* map, rmh_map_info* info);
/* Per-mesh geometry and object ids (host arrays of n_meshes = the map's mesh count; NULL
 * keeps that id).  By default both are the mesh index.  The ids a hit reports follow
 * Embree's: geomID = rtcAttachGeometry order (src/rmagine_embree/src/map/embree/EmbreeScene.cpp:54-71),
 * object id = instID[0] when the hit went through an instance, else geomID
 * (src/rmagine_embree/include/rmagine/simulation/SphereSimulatorEmbree.tcc:175-185). */
int rmh_map_set_mesh_ids(rmh_map* map, const uint32_t* geom_ids, const uint32_t* object_ids, uint32_t n_meshes);
/* Refit: new positions for the map's vertices (device, 3 floats each, same count and faces),
 * triangles, normals and BVH boxes updated in place without rebuilding the tree (the role of
 * OptiX's ALLOW_UPDATE rebuild in OptixScene::commit, src/rmagine_optix/src/map/optix/OptixScene.cpp:296-315,
 * and of EmbreeMesh::apply + commit, src/rmagine_embree/src/map/embree/EmbreeMesh.cpp:297-330).
 * Results equal those of a fresh build; traversal cost grows as vertices move far from the
 * layout the tree was built for.  Synchronous. */
int rmh_map_refit(rmh_map* map, const float* vertices_dev, void* stream);
/* Node layout the ray kernels traverse: wide != 0 builds the 4-wide quantised tree from the
 * binary one (see RMH_BVH_WIDE), wide == 0 frees it.  No counterpart in the reference (its
 * acceleration structures are opaque); a layout the autotuner can choose per map.  A map
 * whose tree cannot be expressed (a single leaf, more than 42 wide levels, coordinates no
 * 8-bit grid holds) stays binary: rmh_map_info.n_wide_nodes then reads 0.  Synchronous; no
 * launch that uses the map may be in flight, and graphs captured with the other layout
 * must be captured again. */
int rmh_map_set_layout(rmh_map* map, int wide, void* stream);

/* Closest points: for n query points (device, 12 B each) the nearest point of the map
 * strictly within max_distance.  Replaces EmbreeMap::closestPoint
 * (src/rmagine_embree/src/map/EmbreeMap.cpp:33-37, EmbreeScene.cpp:292-315, callback
 * EmbreeMesh.cpp:22-131) for a batch.  Every output pointer (device) may be NULL.  Misses:
 * dist FLT_MAX, ids 0xFFFFFFFF, point and normal NaN.  Equal distances: lower global face.
 * Batches of 65,536 queries or more are sorted on the device by the Morton code of their
 * positions and walked in that order when their order is spatially incoherent (random
 * points; not the rows of a scan); the results do not depend on the order.  The sort's
 * scratch (17 B per query) comes from a stream-ordered pool of the library that keeps up to
 * 256 MiB per device between calls; under stream capture the batch is walked as given. */
typedef struct {
  float* dist;
  rmh_vec3* points;
  rmh_vec3* normals;      /* unit face normal of the closest triangle */
  uint32_t* face_ids;
  uint32_t* geom_ids;
  uint32_t* object_ids;
} rmh_closest_outputs;
int rmh_closest_points(const rmh_map* map, const rmh_vec3* queries_dev, uint32_t n, float max_distance,
                       const rmh_closest_outputs* out, void* stream);

/* ---- map files ----------------------------------------------------------------------------
 * Replaces import_embree_map / import_optix_map
 * (src/rmagine_embree/include/rmagine/map/EmbreeMap.hpp:94-117,
 *  src/rmagine_optix/include/rmagine/map/OptixMap.hpp:88-109) without Assimp.  Formats: PLY
 * (ascii, binary little/big endian), OBJ, STL (ascii/binary), COLLADA .dae.  The scene graph
 * is flattened as make_embree_scene + freeze build it (EmbreeScene.cpp:246-290,317-406):
 * vertices transformed by their node's global transform, meshes in final geometry order,
 * with the geom and object ids of that scene.  Host memory only. */
#define RMH_SCENE_POLYGONS_FAN 1u  /* triangulate polygons as fans (default: a polygon keeps
                                      its first three indices, as EmbreeMesh::init does) */
typedef struct rmh_scene rmh_scene;
int rmh_scene_load(const char* path, uint32_t flags, rmh_scene** out);
int rmh_scene_counts(const rmh_scene* scene, uint32_t* n_vertices, uint32_t* n_faces, uint32_t* n_meshes);
/* Copies into host arrays of 3*n_vertices floats, 3*n_faces, n_meshes+1, n_meshes and
 * n_meshes uint32 (any may be NULL). */
int rmh_scene_copy(const rmh_scene* scene, float* vertices, uint32_t* faces, uint32_t* mesh_face_offsets,
                   uint32_t* geom_ids, uint32_t* object_ids);
int rmh_scene_destroy(rmh_scene* scene);

/* O1Dn model of an Ouster sensor from its metadata JSON: replaces
 * o1dn_from_ouster_meta_file (src/rmagine_ouster/src/types/ouster_sensors.cpp:8-80).  Fills
 * width (columns of lidar_mode) and height (beams); with dirs != NULL also the W*H unit
 * directions, row-major (3 floats each, capacity in floats).  The reference's model has
 * origin (0,0,0) and range [0.1, 80]. */
int rmh_ouster_o1dn(const char* path, uint32_t* width, uint32_t* height, float* dirs, size_t capacity);

/* ---- sensors ------------------------------------------------------------------------ */
/* model points at the host struct of `kind` (rmh_spherical_model, rmh_pinhole_model,
 * rmh_o1dn_model, rmh_ondn_model).  The sensor keeps its own device copy of everything
 * it needs (direction tables, O1Dn/OnDn rays). */
int rmh_sensor_create(int kind, const void* model, rmh_sensor** out);
int rmh_sensor_destroy(rmh_sensor* sensor);
/* Traversal order of an O1Dn/OnDn sensor's rays (no counterpart in the reference, which
 * casts them in list order): RMH_RAY_ORDER_SORTED = by direction (a wave's 64 rays cover
 * one patch of the sphere), RMH_RAY_ORDER_LIST = list order (rays already ordered by the
 * caller), RMH_RAY_ORDER_TILES = the model's W x H grid (buffer index row * W + column) in
 * 4 x 16 tiles, as image sensors are cast (for a scanner's ray grid).  Defaults, set by
 * rmh_sensor_create: models of at least 4 rows and 16 columns start in tiles with the
 * classic walk (rmh_sensor_set_walk), other lists sorted with the speculative walk.
 * Results are identical in every order; a tuning choice of the autotuner.  Rewrites the
 * sensor's device arrays in place (batches and graphs that use the sensor stay valid);
 * synchronous; no launch using the sensor may be in flight.  Other values are refused with
 * RMH_ERR_INVALID.  Image sensors (spherical, pinhole) ignore it and report the list order. */
#define RMH_RAY_ORDER_LIST   0
#define RMH_RAY_ORDER_SORTED 1
#define RMH_RAY_ORDER_TILES  2
int rmh_sensor_set_ray_order(rmh_sensor* sensor, int order);
int rmh_sensor_get_ray_order(const rmh_sensor* sensor, int* order);
/* Traversal walk of an O1Dn/OnDn sensor's rays in rmh_simulate and
 * rmh_simulate_visit_stats (no counterpart in the reference, whose ray casts are one
 * library call per ray): speculative != 0 = a lane that reaches a leaf holds it while the
 * wave walks on, and the wave tests leaves once an eighth of its lanes hold one (the
 * default of lists of fewer than 4 rows or 16 columns; fastest for scattered rays such as
 * ICP scans), 0 = the classic walk, in which the wave walks until every lane holds a leaf
 * (the default of larger grids; fastest for a scanner's coherent ray grid).  Results are
 * identical either way; a tuning choice of the autotuner.  A flag read
 * at launch: no device work.  Batches and the pose gradient walk ray lists speculatively
 * whatever the sensor holds.  Image sensors always walk the classic way and report 0. */
int rmh_sensor_set_walk(rmh_sensor* sensor, int speculative);
int rmh_sensor_get_walk(const rmh_sensor* sensor, int* speculative);
int rmh_sensor_dims(const rmh_sensor* sensor, uint32_t* width, uint32_t* height);

/* ---- simulation --------------------------------------------------------------------- */
/* Casts n_poses * W * H rays: for pose p the sensor pose is Tsm = Tbm_dev[p] * Tsb,
 * rays start at Tsm applied to the model origin with direction Tsm.R * d_s, tnear = 0,
 * tfar = range.max, closest hit over two-sided triangles.  Results follow the Embree
 * contract (SphereSimulatorEmbree.tcc:104-249): hit = t >= range.min, range = t (also
 * when t < range.min), points/normals in the sensor frame, normals flipped toward the
 * sensor; misses write hits 0, range.max + 1, NaN points/normals, ids 0xFFFFFFFF.
 * Tsb is a host transform (NULL = identity).  Asynchronous on `stream`. */
int rmh_simulate(const rmh_map* map, const rmh_sensor* sensor,
                 const rmh_transform* Tsb, const rmh_transform* Tbm_dev, uint32_t n_poses,
                 const rmh_outputs* out, void* stream);
/* Diagnostic, no counterpart in the reference: the traversal of rmh_simulate's kernel over
 * the same rays (binary node layout, same speculation mode), counting instead of writing.
 * counts (host, 6 entries, summed over all rays): [0] node visits read with vector loads,
 * [1] node visits read through the scalar cache (the wave at one node), [2] leaves,
 * [3] triangle tests, [4] wave iterations of the node loop, [5] of them wave-uniform.
 * Synchronous. */
int rmh_simulate_visit_stats(const rmh_map* map, const rmh_sensor* sensor, const rmh_transform* Tsb,
                             const rmh_transform* Tbm_dev, uint32_t n_poses, uint64_t* counts, void* stream);

/* ---- batched heterogeneous simulation ---------------------------------------------------
 * No counterpart in the reference (it launches one sensor model per optixLaunch,
 * src/rmagine_optix/src/simulation/SphereSimulatorOptix.cpp:64-81; the closest piece is the
 * model union of src/rmagine_optix/include/rmagine/simulation/optix/sim_program_data.h:32-44).
 * A batch is a device table of sensors (any mix of models, poses and ray counts) cast in one
 * persistent launch (rmh_batch_params.launches = 2 splits it into image sensors, then O1Dn/OnDn
 * ray lists); every entry writes its own rmh_outputs exactly as rmh_simulate would. */
typedef struct {
  const rmh_sensor*    sensor;
  rmh_transform        Tsb;        /* sensor -> base (host value)               */
  const rmh_transform* Tbm_dev;    /* n_poses base -> map poses (device)        */
  uint32_t             n_poses;
  uint32_t             _pad;
  rmh_outputs          out;        /* device buffers, n_poses*W*H entries each  */
} rmh_batch_entry;

typedef struct {
  uint32_t chunk_rays;    /* rays per work grab, multiple of 64 (0 = keep / default 256)  */
  uint32_t waves_per_cu;  /* persistent waves per compute unit (0 = keep / default 32)   */
  uint32_t tile_rows;     /* image rows per 64-ray tile of spherical/pinhole sensors:
                             1, 2, 4, 8, 16, 32 or 64, capped by the image height
                             (0 = keep / default 4)                                      */
  uint32_t lds_stack;     /* traversal stack entries per lane kept in LDS: 8, 10 or 16
                             (8 B each; 0 = keep / default 10); deeper entries go to
                             scratch                                                     */
  uint32_t launches;      /* persistent launches per rmh_batch_simulate: 1 = every entry in
                             one launch (0 = keep / default), 2 = image sensors, then ray
                             lists, each kernel with its own walk                       */
} rmh_batch_params;

typedef struct rmh_batch rmh_batch;

/* Builds the device table (the pointers in the entries must stay valid while the batch is
 * used).  Synchronous. */
int rmh_batch_create(const rmh_map* map, const rmh_batch_entry* entries, uint32_t n_entries,
                     const rmh_batch_params* params, rmh_batch** out);
int rmh_batch_set_params(rmh_batch* batch, const rmh_batch_params* params);
int rmh_batch_get_info(const rmh_batch* batch, uint64_t* total_rays, uint32_t* total_chunks,
                       rmh_batch_params* params);
/* One counter-reset kernel + one persistent kernel on `stream` (per group when launches = 2,
 * in stream order); capturable into a graph.  Launches of one batch must not overlap (they
 * share its work counters). */
int rmh_batch_simulate(rmh_batch* batch, void* stream);
int rmh_batch_destroy(rmh_batch* batch);

/* ---- range noise -------------------------------------------------------------------------
 * Replaces GaussianNoiseCuda / RelGaussianNoiseCuda / UniformDustNoiseCuda::apply
 * (src/rmagine_cuda/src/noise/GaussianNoiseCuda.cu:7-48, RelGaussianNoiseCuda.cu:7-54,
 *  UniformDustNoiseCuda.cu:7-69), in place on a device range buffer.  The random numbers are
 * counter-based (Philox4x32-10 of element index and `epoch` under `seed`): no per-element
 * state; pass a new epoch for new numbers (the wrappers count calls). */
#define RMH_NOISE_GAUSSIAN 0u      /* p0 = mean, p1 = stddev                              */
#define RMH_NOISE_REL_GAUSSIAN 1u  /* p0 = mean, p1 = stddev, p2 = range exponent         */
#define RMH_NOISE_UNIFORM_DUST 2u  /* p0 = hit probability per metre, p1 = return prob.   */
typedef struct {
  uint32_t kind;
  float p0, p1, p2;
  float max_range;  /* Gaussian kinds: ranges above are left alone; dust: range cap  */
  uint32_t seed;
} rmh_noise_params;
int rmh_noise_apply(const rmh_noise_params* params, float* ranges_dev, uint64_t n, uint32_t epoch, void* stream);

/* ---- graphs: record a sequence of rmh_* launches on a stream once, replay it ------------
 * No counterpart in the reference.  Capture needs a non-default stream; nothing captured
 * may allocate or synchronise (pass workspaces to the reductions). */
typedef struct rmh_graph rmh_graph;
int rmh_graph_begin_capture(void* stream);
int rmh_graph_end_capture(void* stream, rmh_graph** out);
int rmh_graph_launch(rmh_graph* graph, void* stream);
int rmh_graph_destroy(rmh_graph* graph);

/* ---- pose adjoint ------------------------------------------------------------------------
 * No counterpart in the reference (closest: jacobian_and_residual_p2l,
 * src/rmagine_core/include/rmagine/math/optimization.h:113-137; se3/so3 Jacobians,
 * src/rmagine_core/include/rmagine/math/lie.tcc:252-277).
 * For every pose: forward closest-hit ranges r (as rmh_simulate), Huber(huber_delta) loss of
 * r - target over rays that hit (r >= range.min) and whose target is valid (target_mask != 0,
 * or range.min <= target <= range.max when target_mask is NULL), and its gradient with
 * respect to a left perturbation of the pose in the map frame (t + tau, Exp(omega) R).
 * target: W*H floats shared by all poses (per_pose_targets = 0) or n_poses*W*H.
 * out_dev: 8 doubles per pose {loss, dL/dtau xyz, dL/domega xyz, rays used}.
 * workspace: rmh_pose_grad_workspace_size bytes of device memory, or NULL (then the call
 * allocates and synchronises internally). */
int rmh_pose_grad_workspace_size(const rmh_map* map, const rmh_sensor* sensor, uint32_t n_poses, size_t* bytes);
int rmh_pose_grad(const rmh_map* map, const rmh_sensor* sensor, const rmh_transform* Tsb,
                  const rmh_transform* Tbm_dev, uint32_t n_poses, const float* target_dev,
                  const uint8_t* target_mask_dev, int per_pose_targets, float huber_delta,
                  double* out_dev, void* workspace, size_t workspace_bytes, void* stream);

/* ---- math layer (rmagine_cuda's math, src/rmagine_cuda/src/math/) --------------------- */

/* C[i] = A[nA == 1 ? 0 : i] * B[nB == 1 ? 0 : i], i < n  (multNxN / multNx1 / mult1xN of
 * Transforms, src/rmagine_cuda/src/math/memory_math.cu:16-73,471-1024).  Device arrays. */
int rmh_transform_mul(const rmh_transform* A_dev, uint32_t nA, const rmh_transform* B_dev, uint32_t nB,
                      rmh_transform* C_dev, uint32_t n, void* stream);
/* C[i] = A[i]^-1  (invert, memory_math.cu:121-160, Transform.tcc:24-31) */
int rmh_transform_inv(const rmh_transform* A_dev, rmh_transform* C_dev, uint32_t n, void* stream);
/* out[i] = T[nT == 1 ? 0 : i] * p[i]  (multNxN / mult1xN Transform x Vector) */
int rmh_transform_points(const rmh_transform* T_dev, uint32_t nT, const rmh_vec3* p_dev, rmh_vec3* out_dev,
                         uint32_t n, void* stream);
/* normalizeInplace of the rotation quaternions (memory_math.cu:402-413) */
int rmh_quat_normalize(rmh_transform* T_dev, uint32_t n, void* stream);
/* T[i] = identity (setIdentity_kernel, memory_math.cu:1342-1352); a kernel, never a memset,
 * so it can be captured into a graph */
int rmh_transform_set_identity(rmh_transform* T_dev, uint32_t n, void* stream);
/* pack (src/rmagine_cuda/src/math/memory_math.cu:1258-1279, kernels :313-339): T[i].t = t[i]
 * and T[i].R from R_dev[i], a rotation matrix of 9 floats in column-major order (Rmagine's
 * Matrix3x3 layout) converted with the reference's FP32 matrix -> quaternion steps
 * (MatrixOps.tcc:1492-1529, no normalisation), or from q_dev[i]; T[i].stamp is left as it
 * was.  Device arrays of n entries. */
int rmh_transform_pack_matrix(const float* R_dev, const rmh_vec3* t_dev, rmh_transform* T_dev, uint32_t n,
                              void* stream);
int rmh_transform_pack_quat(const rmh_quat* q_dev, const rmh_vec3* t_dev, rmh_transform* T_dev, uint32_t n,
                            void* stream);

/* Workspace (device bytes) that rmh_vec3_moments / rmh_statistics / rmh_linear_system_p2l
 * need for this shape when a segment is split over several blocks (one size serves all
 * three); 0 when none is needed. */
int rmh_reduction_workspace_size(uint64_t segment_size, uint32_t n_segments, size_t* bytes);

/* Segmented FP64 moments of n_segments consecutive segments of segment_size vectors:
 * sum_a (3/segment), mean_a (3/segment), cov = (1/n) sum b a^T (9/segment, column-major,
 * b = a when b_dev is NULL) and count, over the elements whose mask is nonzero (mask NULL:
 * all).  Replaces sum/mean/cov (memory_math.cu:415-469,1403-1549) and
 * sumBatched/covBatched (math_batched.cu:30-546).  Outputs are nullable device arrays.
 * workspace: device memory of rmh_reduction_workspace_size bytes, or NULL (then the call
 * allocates and synchronises internally). */
int rmh_vec3_moments(const rmh_vec3* a_dev, const rmh_vec3* b_dev, const uint8_t* mask_dev,
                     uint64_t segment_size, uint32_t n_segments, double* sum_a_dev, double* mean_a_dev,
                     double* cov_dev, uint32_t* count_dev, void* workspace, size_t workspace_bytes,
                     void* stream);

/* A point cloud view (PointCloudView_, src/rmagine_core/include/rmagine/types/PointCloud.hpp:62-70)
 * over device arrays; mask, normals and ids are nullable.  Element i of segment s is at
 * s * segment_stride + i (segment_stride 0: every segment shares the same elements). */
typedef struct {
  const rmh_vec3* points;
  const uint8_t*  mask;
  const rmh_vec3* normals;
  const uint32_t* ids;
  uint64_t        segment_stride;
} rmh_cloud;

/* rmagine::UmeyamaReductionConstraints (UmeyamaReductionConstraints.hpp:46-59) */
typedef struct { float max_dist; uint32_t dataset_id; uint32_t model_id; } rmh_umeyama_constraints;

/* FP64 CrossStatistics (CrossStatistics.hpp:50-60): means, covariance
 * (1/n) sum (m - m_mean)(d - d_mean)^T column-major, count.  128 bytes. */
typedef struct {
  double   dataset_mean[3];
  double   model_mean[3];
  double   covariance[9];
  uint32_t n_meas;
  uint32_t _pad;
} rmh_cross_stats;

enum { RMH_STATS_P2P = 0, RMH_STATS_P2L = 1 };

/* Correspondence statistics per segment (statistics_p2p / statistics_p2l,
 * src/rmagine_cuda/src/math/statistics.cu:58-160,190-276,374-420; CPU
 * src/rmagine_core/src/math/statistics.cpp:20-126), in FP64: element i passes the mask
 * and id filters, D = pre_transform[s] * dataset point, then
 *   p2p: M = model point, kept if |M - D| < max_dist;
 *   p2l: M = D + N ((I - D) . N) with I, N the model point and normal, kept if
 *        |(I - D) . N| < max_dist.
 * pre_transforms_dev holds n_pre = 1 (shared) or n_segments transforms. */
int rmh_statistics(int kind, const rmh_transform* pre_transforms_dev, uint32_t n_pre,
                   const rmh_cloud* dataset, const rmh_cloud* model, uint64_t segment_size,
                   uint32_t n_segments, const rmh_umeyama_constraints* params,
                   rmh_cross_stats* stats_dev, void* workspace, size_t workspace_bytes, void* stream);

/* rmh_statistics without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc) and pass no workspace.  Same arguments, kernels and results as
 * rmh_statistics; the per-block partial sums go to a fixed 2 MiB device array of the library
 * (one per GPU, held until the process ends), which covers every shape rmh_statistics
 * accepts (larger partial sets would be refused with RMH_ERR_UNSUPPORTED).  Calls that
 * overlap on one GPU share that array and race: serialise them (one stream). */
int rmh_statistics_noalloc(int kind, const rmh_transform* pre_transforms_dev, uint32_t n_pre,
                           const rmh_cloud* dataset, const rmh_cloud* model, uint64_t segment_size,
                           uint32_t n_segments, const rmh_umeyama_constraints* params,
                           rmh_cross_stats* stats_dev, void* stream);

/* rmagine::AABB (src/rmagine_core/include/rmagine/math/types/AABB.hpp:50-53), 24 bytes */
typedef struct { rmh_vec3 min, max; } rmh_aabb;

/* Objectwise point-to-plane statistics (statistics_objectwise_p2l,
 * src/rmagine_cuda/src/math/statistics.cu:278-371,422-470; declared statistics.cuh:139-158),
 * FP64: one rmh_cross_stats per object o < n_objects over image-shaped scans of n_points
 * elements with row stride `width` (height is accepted as in the reference and unused).
 * bboxes_dev[o] is the object's pixel window (min.x/max.x columns, min.y/max.y rows, max
 * exclusive; the floats are converted to unsigned like the reference's assignments, NaN and
 * negatives giving 0); window element k (row-major, k < w*h) is scan element
 * (min.y + k / w) * width + min.x + k % w, used while below n_points.  The reference's kernel
 * runs k up to w*h rounded up to a multiple of 512, into the rows below the window; that
 * element set is the default, RMH_OBJECTWISE_WINDOW_ONLY stops at w*h.  An inverted window
 * is empty.  Each element passes the mask filters, the id filters and the plane-distance
 * test of params_dev[o] under pre_transforms_dev[o], as in rmh_statistics (p2l).  One
 * workgroup per object; no workspace. */
#define RMH_OBJECTWISE_WINDOW_ONLY 1u
int rmh_statistics_objectwise_p2l(const rmh_transform* pre_transforms_dev, const rmh_cloud* dataset,
                                  const rmh_cloud* model, uint32_t width, uint32_t height, uint32_t n_points,
                                  const rmh_umeyama_constraints* params_dev, const rmh_aabb* bboxes_dev,
                                  uint32_t n_objects, uint32_t flags, rmh_cross_stats* stats_dev, void* stream);

/* Point-to-plane Gauss-Newton system per segment: build_linear_system_p2l
 * (src/rmagine_core/include/rmagine/math/optimization.h:170-200) with
 * jacobian_and_residual_p2l (:113-137), in FP64.  With Pd = pre_transform * dataset point,
 * Pm, Nm the model point and normal: r = (Pd - Pm) . Nm, J = [Pd x Nm, Nm] (the derivative
 * of r for the left perturbation Pd -> Exp(omega) Pd + tau, unknowns (omega, tau)), weight
 * w = GM_weight(s, r) = s^2 / (s + r)^2 with s = kernel_scale (the reference passes the signed
 * residual and s = 5; kernel_scale <= 0 gives w = 1).  Sums J^T w J, J^T w r = the gradient
 * of (1/2) sum w r^2 with w held fixed, and sum w r^2 over the correspondences that pass the
 * mask/id filters and |r| < max_dist (rmh_statistics' filters).  Segments, pre-transforms and
 * the workspace as in rmh_statistics. */
typedef struct {
  double   JTwJ[36];  /* row-major 6x6 (symmetric) */
  double   JTwr[6];
  double   wr2;       /* sum of w r^2 */
  uint32_t n_meas;
  uint32_t _pad;
} rmh_p2l_system;
int rmh_linear_system_p2l(const rmh_transform* pre_transforms_dev, uint32_t n_pre, const rmh_cloud* dataset,
                          const rmh_cloud* model, uint64_t segment_size, uint32_t n_segments,
                          const rmh_umeyama_constraints* params, double kernel_scale, rmh_p2l_system* out_dev,
                          void* workspace, size_t workspace_bytes, void* stream);

/* Umeyama per statistics entry (umeyama_transform, src/rmagine_core/src/math/optimization.cpp:12-55;
 * src/rmagine_cuda/src/math/linalg.cu:1150-1190) in FP64: out_dev receives FP32 transforms,
 * out64_dev (nullable) 7 doubles per entry (qx, qy, qz, qw, tx, ty, tz). */
int rmh_umeyama(const rmh_cross_stats* stats_dev, uint32_t n, rmh_transform* out_dev, double* out64_dev,
                void* stream);
/* The same FP64 Umeyama on the host (host arrays): replaces the CPU umeyama_transform
 * (src/rmagine_core/src/math/optimization.cpp:12-55) with the device's algorithm. */
int rmh_umeyama_host(const rmh_cross_stats* stats, uint32_t n, rmh_transform* out, double* out64);

/* FP64 statistics -> Rmagine's CrossStatistics_<float> (CrossStatistics.hpp:50-60: dataset
 * mean, model mean, column-major covariance as 15 floats, then n_meas; 64 bytes each), device
 * arrays; and FP32 (dataset means, model means, column-major covariances, counts or NULL for
 * n_meas = 1) -> FP64 statistics, for the batched umeyama_transform(ds, ms, Cs, n_meas) of
 * src/rmagine_cuda/include/rmagine/math/memory_math.cuh:590-618. */
int rmh_cross_stats_to_f32(const rmh_cross_stats* in_dev, void* out_dev, uint32_t n, void* stream);
int rmh_cross_stats_pack(const rmh_vec3* dataset_means_dev, const rmh_vec3* model_means_dev,
                         const float* covariances_dev, const uint32_t* n_meas_dev, rmh_cross_stats* out_dev,
                         uint32_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RMH_H */
