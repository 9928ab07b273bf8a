/*
 * kdop_amd.h -- C ABI of the MI355X kDOP BVH collision library.
 *
 * Drop-in boundary for the KDOP subsystem of diku-dk/bvh-tvcg18 (the PROX
 * simulator).  The reference exposes a tag-dispatched C++ template API
 * (kdop::make_tree / refit_tree / tandem_traversal, narrow::update_kdop_bvh);
 * each entry point below names the reference call it replaces.  All entry
 * points are synchronous (they return once the device work they start has
 * finished; refit can opt out, KDOP_OPT_ASYNC_REFIT) and return 0 (KDOP_OK) or
 * a negative KDOP_ERR_* code; the message of the last failure is
 * kdop_last_error(ctx).  One context per host thread.
 *
 * Pointers are host pointers unless a parameter says "device".
 */
#ifndef KDOP_AMD_H
#define KDOP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KDOP_OK 0
#define KDOP_ERR_ARG (-1)      /* invalid argument (bad K, empty mesh, unknown body id, ...) */
#define KDOP_ERR_HIP (-2)      /* HIP runtime error (no device, launch failure, ...)        */
#define KDOP_ERR_NOMEM (-3)    /* device allocation failed                                  */
#define KDOP_ERR_STATE (-4)    /* call out of order (e.g. kdop_narrow before a traversal)   */
#define KDOP_ERR_CAPACITY (-5) /* caller buffer too small; the required size is reported    */
#define KDOP_ERR_INTERNAL (-6) /* an internal consistency check failed (a library bug)      */

typedef struct kdop_ctx kdop_ctx;

/* Traversal strategies (kdop_config.traversal). */
#define KDOP_TRAVERSAL_AUTO 0        /* persistent for K >= 12, >= 64 body pairs or fewer  */
                                     /* than 400k tets in the call; level-synchronous      */
                                     /* otherwise (kdop_autotune refines)                  */
#define KDOP_TRAVERSAL_LEVELS 1      /* level-synchronous frontier kernels                 */
#define KDOP_TRAVERSAL_PERSISTENT 2  /* persistent rounds: LDS stacks per workgroup, guided */
                                     /* frontier chunks, budgeted rounds re-spreading work  */

typedef struct kdop_config {
  int32_t K;             /* slab count of the kDOP: 6, 8, 12, 14, 18, 20 or 26 (template K)    */
  int32_t device;        /* HIP device ordinal                                                  */
  uint64_t chunk_bytes;  /* reference make_tree mem_bytes (narrow_params.h:54 default 8000)     */
  int32_t traversal;     /* KDOP_TRAVERSAL_*                                                    */
  int32_t queue_depth;   /* persistent kernel: LDS stack entries per wave (0 = autotune)       */
  int32_t waves_per_cu;  /* persistent kernel: resident waves per CU (0 = autotune)            */
  int32_t ramp_target;   /* persistent kernel: frontier size that ends the breadth-first ramp   */
  int32_t stack_keep;    /* persistent kernel: stack entries a wave keeps once the frontier is  */
                         /*   exhausted; deeper stacks go to the next round (0 = default: no  */
                         /*   hand-off, KDOP_OPT_STACK_KEEP)                                  */
  int32_t seed_budget;   /* chunk-root pairs tested to seed one call (0 = by the call's pairs)  */
  int32_t reserved[2];
} kdop_config;

/* Counters of one traversal (+ narrow phase) call. */
typedef struct kdop_counters {
  uint64_t tests;            /* node-pair DOP tests the device executed (incl. the root test)   */
  uint64_t reference_tests;  /* tests the reference CPU traversal makes for the same input:      */
                             /*   1 + C_A*C_B + descent_tests when the roots overlap, else 1     */
  uint64_t descent_tests;    /* tests below branch-root pairs (identical on both sides)          */
  uint64_t branch_overlaps;  /* overlapping branch-root pairs                                    */
  uint64_t leaf_pairs;       /* overlapping leaf (tet) pairs, P1                                  */
  uint64_t surface_pairs;    /* P1 pairs whose tets both have a surface face, P2                  */
  uint64_t contacts;         /* contact points reported by the narrow phase                      */
  uint64_t narrow_pairs;     /* P2 pairs for which the narrow phase generated contacts           */
  uint64_t branch_pairs;     /* C_A*C_B summed over body pairs whose roots overlap               */
  uint64_t overflow;         /* nonzero if an internal buffer had to grow (the call was re-run)  */
} kdop_counters;

/* One contact point (geometry::ContactsCallback arguments, contacts/geometry_contacts_callback.h:7-24). */
typedef struct kdop_contact {
  uint32_t pair;      /* index into the body-pair list of the traversal call                 */
  uint32_t tet_a;     /* tetrahedron of body A                                               */
  uint32_t tet_b;     /* tetrahedron of body B                                               */
  uint32_t ordinal;   /* position among this tet pair's reported contacts (reference order)  */
  float p[3];         /* contact point, world coordinates                                    */
  float n[3];         /* unit contact normal (a face normal of tet_a)                        */
  float depth;        /* penetration measure, negative when overlapping                      */
} kdop_contact;

/* One sphere-tetrahedron contact of kdop_single_traversal (the arguments of
 * the ContactsCallback call in contacts/geometry_contacts_sphere_tetrahedron.h). */
typedef struct kdop_sphere_contact {
  uint32_t sphere;    /* index into the sphere list of the call        */
  uint32_t tet;       /* tetrahedron of the body                       */
  float p[3];         /* contact point                                 */
  float n[3];         /* unit normal (sphere to tetrahedron unless flip) */
  float depth;        /* negative when penetrating                     */
} kdop_sphere_contact;

/* One contact as PROX holds it (prox::ContactPoint, SIMULATION/PROX/PROX/
 * include/prox_contact_point.h:15-63): the two bodies, point, unit normal and
 * penetration measure.  Bodies are numbered by the caller (PROX compares body
 * pointers; a body's index in the caller's body array orders them the same way
 * when the bodies sit in one array). */
typedef struct kdop_prox_contact {
  uint32_t body_i;
  uint32_t body_j;
  float p[3];
  float n[3];
  float depth;
} kdop_prox_contact;

/* Library defaults: K=8, chunk_bytes=8000, device 0, autotuned traversal. */
void kdop_default_config(kdop_config* cfg);

/* Create / destroy a context (owns a HIP stream on cfg->device and all device
 * buffers).  Replaces the DIKUCL singletons (FOUNDATION/DIKUCL dikucl_*_manager.hpp). */
int kdop_create(kdop_ctx** ctx, const kdop_config* cfg);
void kdop_destroy(kdop_ctx* ctx);
const char* kdop_last_error(const kdop_ctx* ctx);

/* Upload a tetrahedral mesh (material coordinates X0/Y0/Z0, tets as 4 vertex
 * indices i,j,k,m, surface bits: bit v set when the face opposite vertex v is on
 * the surface) and build its chunked kDOP tree on the device, refitted to the
 * material coordinates, with the context's chunk_bytes as mem_bytes.
 * Replaces kdop::make_tree<V,K,T>(mem_bytes, mesh, X, Y, Z, sequential)
 * (SIMULATION/KDOP/KDOP/include/kdop_make_tree.h:190-288) as called from
 * narrow::make_kdop_bvh (SIMULATION/NARROW/NARROW/include/narrow_object.h:96-134). */
int kdop_add_body(kdop_ctx* ctx, uint32_t nv, const float* X0, const float* Y0, const float* Z0,
                  uint32_t nt, const uint32_t* tets, const uint8_t* surface, uint32_t* body_id);

/* The same with the tree shape of this body's own mem_bytes, the first
 * argument of kdop::make_tree (kdop_make_tree.h:190-229): the chunk size
 * follows from how many kdop::Node<float,K> fit in mem_bytes / 2, so any
 * value the reference accepts is accepted here, including chunks of more
 * than 32 leaves (narrow_chunk_bytes from the PROX configuration,
 * simulators_prox_parameters.cpp:54,394) and SIZE_MAX, one chunk of every
 * tet (narrow_object.h:104-111).  Bodies of one context may differ in
 * mem_bytes.  Chunks wider than 32 leaves get their topology from the host
 * and a refit that unions one tree height per launch; the results are the
 * same bits either way. */
int kdop_add_body_ex(kdop_ctx* ctx, uint64_t mem_bytes, uint32_t nv, const float* X0, const float* Y0, const float* Z0,
                     uint32_t nt, const uint32_t* tets, const uint8_t* surface, uint32_t* body_id);

/* Tree shape of a body: out[0..4] = {L leaves per chunk, C branch chunks, H levels,
 * device node count, depth (longest root-to-leaf path)}. */
int kdop_body_shape(kdop_ctx* ctx, uint32_t body, uint64_t* out);

/* Rigid transform + refit of n bodies: r = rotate(q, r0) + p per vertex, then
 * the leaf and union passes.  pos: 3n floats, quat: 4n floats (x, y, z, w).
 * Replaces narrow::update_kdop_bvh(work_pool, sequential)
 * (SIMULATION/NARROW/NARROW/include/narrow_update_kdop_bvh.h:126-176), whose
 * per-body refit is kdop::refit_tree (SIMULATION/KDOP/KDOP/include/kdop_refit_tree.h:119-162). */
int kdop_refit(kdop_ctx* ctx, uint32_t n, const uint32_t* bodies, const float* pos, const float* quat);

/* Same with the poses already in device memory (pos, quat: device pointers). */
int kdop_refit_device(kdop_ctx* ctx, uint32_t n, const uint32_t* bodies, const float* pos_dev, const float* quat_dev);

/* Flags of kdop_traverse. */
#define KDOP_NARROW 1u         /* also run the tet-tet narrow phase on the surface pairs    */
#define KDOP_KEEP_ALL_PAIRS 2u /* keep the full leaf-pair list P1 (else only P2 is kept)    */

/* Tandem traversal of npairs body pairs (body_pairs: 2*npairs ids, A then B)
 * in one device pass; with KDOP_NARROW the MOST_OPPOSING_SURFACES contact
 * generator runs on every surface pair.  Results stay in device memory until
 * the next traversal (kdop_get_pairs / kdop_get_contacts).
 * Replaces kdop::tandem_traversal(std::vector<TestPair>&, sequential|dikucl)
 * (SIMULATION/KDOP/KDOP/include/kdop_tandem_traversal.h:198-226,
 *  SIMULATION/KDOP/KDOP/include/cl/kdop_cl_tandem_traversal.h:622-626) as called from
 * narrow::details::dispatch_tetramesh_tetramesh (narrow_tetramesh_tetramesh.h:20-93). */
int kdop_traverse(kdop_ctx* ctx, uint32_t npairs, const uint32_t* body_pairs, uint32_t flags, kdop_counters* out);

/* Copy the leaf pairs of the last traversal: which = 0 for P1 (needs
 * KDOP_KEEP_ALL_PAIRS), 1 for P2.  dst receives 3 uint32 per pair (pair index,
 * tet_a, tet_b) in device order (unsorted).  *n receives the pair count; returns
 * KDOP_ERR_CAPACITY (nothing copied) when cap < *n.  dst_is_device: dst is a
 * device pointer. */
int kdop_get_pairs(kdop_ctx* ctx, int which, uint32_t* dst, uint64_t cap, int dst_is_device, uint64_t* n);

/* Copy the contacts of the last traversal, in device order unless
 * kdop_sort_contacts or kdop_reduce_contacts ran after it. */
int kdop_get_contacts(kdop_ctx* ctx, kdop_contact* dst, uint64_t cap, int dst_is_device, uint64_t* n);

/* Orders of kdop_sort_contacts. */
#define KDOP_ORDER_SEQUENTIAL 0 /* the callback order of kdop::tandem_traversal(..., sequential) */
#define KDOP_ORDER_OPENCL 1     /* (pair, p, n, depth), float comparisons: the OpenCL path's report order */

/* Sort the contacts of the last KDOP_NARROW traversal on the device, in place,
 * so that kdop_get_contacts returns them in that order.  KDOP_ORDER_SEQUENTIAL
 * is the order the sequential traversal reports them through each pair's
 * ContactsCallback: body pair, branch chunk pairs row-major
 * (SIMULATION/KDOP/KDOP/include/kdop_tandem_traversal.h:160-196), the
 * depth-first descent inside a chunk pair (:18-156), then the contact
 * generator's order (the ordinal); KDOP_ORDER_OPENCL replaces the
 * host qsort of the OpenCL path (SIMULATION/KDOP/KDOP/include/cl/
 * kdop_cl_tandem_traversal.h:61-104 compare_kernel_contact_points, applied at
 * :1772-1788), with the source position breaking ties that qsort leaves open. */
int kdop_sort_contacts(kdop_ctx* ctx, int order);

/* PROX contact reduction, on the device, of the contacts of the last
 * KDOP_NARROW traversal, in place: body i and body j of a contact are its
 * pair's bodies A and B (the ContactCallbackFunctor of
 * SIMULATION/PROX/PROX/include/prox_collision_detection.h:38-110).  Replaces
 * the std::sort + std::unique of prox_collision_detection.h:340-345 over
 * prox::ContactPoint: sorted by (body i, body j, p.y, p.x, p.z, n.y, n.x,
 * n.z, depth) with float comparisons (prox_contact_point.h:140-190), then a
 * contact is dropped when it equals the last one kept, equal meaning the same
 * bodies and every compared float within tiny::working_precision<float>()
 * (prox_contact_point.h:91-138, tiny_precision.h:21-24).  Points equal in every
 * compared field (which std::sort leaves in no particular order) keep their
 * traversal order.  *n receives the count left; kdop_get_contacts then
 * returns the reduced contacts in that sorted order (each keeps its pair,
 * tets and ordinal). */
int kdop_reduce_contacts(kdop_ctx* ctx, uint64_t* n);

/* The same reduction over a caller's contact array (in: n contacts; out: room
 * for n; both device pointers when is_device, else host; out must not alias
 * in).  *n_out receives the count written.  For contacts of several narrow
 * phases gathered into one array, as PROX gathers them. */
int kdop_reduce_prox_contacts(kdop_ctx* ctx, uint64_t n, const kdop_prox_contact* in, kdop_prox_contact* out, int is_device,
                              uint64_t* n_out);

/* The same reduction with no allocation: in_dev and out_dev are device
 * arrays of n records (not aliased), and the sort keys, flags and scans live
 * in one fixed array of the library per GPU, held until the process ends,
 * sized for at most KDOP_NOALLOC_MAX_CONTACTS records (about 23 MB); a larger
 * n returns KDOP_ERR_CAPACITY and nothing runs.  The call allocates no device
 * or host memory (no hipMalloc, hipFree or allocating library routine), so
 * callers that cannot allocate, such as tight serving loops, can use it;
 * like every call it completes before returning.  Every call on one GPU uses
 * that GPU's one array: calls that overlap on the same GPU (from several
 * contexts or threads) race.  Results equal kdop_reduce_prox_contacts' with
 * is_device = 1.  No counterpart in the reference; it exists for the callers
 * that cannot allocate. */
#define KDOP_NOALLOC_MAX_CONTACTS 262144u
int kdop_reduce_prox_contacts_noalloc(kdop_ctx* ctx, uint64_t n, const kdop_prox_contact* in_dev, kdop_prox_contact* out_dev,
                                      uint64_t* n_out);

/* Morton reorder of a tetrahedral mesh on the device (host arrays in and
 * out; outputs must not alias inputs): tets sorted by the Morton code of their
 * centroid, vertices renumbered in first-touch order, vertices no tet names
 * appended in input order.  Replaces kdop::mesh_reorder (SIMULATION/KDOP/KDOP/
 * include/kdop_mesh_reorder.h:172-353, called from narrow_shape_types.h:172-180)
 * with equal codes kept in input order (the reference's std::sort leaves them in
 * an implementation-defined order); same result as the host kdop_mesh_reorder
 * of kdop_meshgen.h. */
int kdop_mesh_reorder_gpu(kdop_ctx* ctx, uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt,
                          const uint32_t* tets, float* Xo, float* Yo, float* Zo, uint32_t* tets_o);

/* Cast n rays at one body in its last refitted pose (origins and directions
 * 3 floats each, a direction need not be unit: it is normalised as
 * geometry::make_ray does; is_device 0: host arrays, 1: device arrays of the
 * context's GPU, read and written in place, 4-byte aligned).  Per ray: the nearest hit on a front-facing
 * surface face (hit_points, 3 floats), its ray length, and hits[i] = 1; a
 * miss gives the point (0, 0, 0), length FLT_MAX and 0.  Replaces
 * kdop::raycast(ray, tree, mesh, X, Y, Z, surface_map, hit_point, length)
 * (SIMULATION/KDOP/KDOP/include/kdop_raycast.h:102-155) with the same results:
 * the same tests in the same order at and below the branch-chunk roots, with
 * the super levels culling only chunks whose root the reference's test would
 * reject or prune. */
int kdop_raycast(kdop_ctx* ctx, uint32_t body, uint32_t n, const float* origins, const float* directions, float* hit_points,
                 float* lengths, uint8_t* hits, int is_device);

/* Spheres against one body in its last refitted pose (spheres holds x, y, z,
 * radius per sphere; is_device 0: spheres and out are host arrays, 1: device
 * arrays of the context's GPU, spheres 16-byte aligned).  Contacts go to out grouped by sphere in
 * sphere order, each sphere's in the order the reference reports them through
 * its ContactsCallback; *n receives the contact count, and KDOP_ERR_CAPACITY
 * (nothing written) means cap < *n.  The contacts are computed in full on
 * every call, so a cap that is large enough the first time saves a second
 * traversal.  flip: the reference's should_flip.
 * Replaces kdop::single_traversal(sphere, tree, mesh, X, Y, Z, surface_map,
 * callback, should_flip) (SIMULATION/KDOP/KDOP/include/kdop_single_traversal.h:
 * 123-161) for n spheres at once, with the same contacts in the same order. */
int kdop_single_traversal(kdop_ctx* ctx, uint32_t body, uint32_t nspheres, const float* spheres, int flip,
                          kdop_sphere_contact* out, uint64_t cap, uint64_t* n, int is_device);

/* World coordinates of a body after the last refit (3 arrays of nv floats). */
int kdop_get_world(kdop_ctx* ctx, uint32_t body, float* X, float* Y, float* Z);

/* Export a body's device tree: node count in *n; when the buffers are non-null
 * (cap >= *n), per node: vol (K floats lo0,hi0,lo1,...), range (2 uint32: first
 * and last tet covered), kind (1 byte: 0 super-internal, 1 branch-internal,
 * 2 tet leaf), and the root volume (K floats).  For parity tests. */
int kdop_export_tree(kdop_ctx* ctx, uint32_t body, uint64_t cap, float* vol, uint32_t* range, uint8_t* kind,
                     float* root, uint64_t* n);

/* Adjoint of the narrow-phase penetration measure.  Phi(pair) = sum of the
 * depths of the contacts reported for that body pair by the last KDOP_NARROW
 * traversal, with every discrete choice of that traversal frozen (contact
 * set, normal face, provenance of the extreme candidates), evaluated in FP64
 * from the material coordinates and the poses of the last refit, as a smooth
 * function of the poses.  For each body in bodies[] (host array) writes
 * dL/dp (grad_pos: 3 floats per body) and dL/dq (grad_quat: 4 floats, x y z w)
 * with L = sum over pairs of dphi[pair] * Phi(pair) (dphi: device, one float
 * per body pair of the traversal, or NULL for all ones); a body in several
 * pairs gets the sum.  phi (device, one float per pair, may be NULL) receives
 * Phi.  The quaternion gradient is that of the literal r = Im(q r0 q*) + p (q
 * not renormalised).  grad_pos / grad_quat / phi are device pointers.
 * No counterpart in the reference (it has no differentiable narrow phase). */
int kdop_narrow_adjoint(kdop_ctx* ctx, uint32_t nbodies, const uint32_t* bodies, const float* dphi_dev,
                        float* phi_dev, float* grad_pos_dev, float* grad_quat_dev);

/* Device wall time (milliseconds) of the phases of the last traversal call,
 * measured with HIP events on the context stream: out[0] traversal kernels,
 * out[1] narrow phase, out[2] whole call.  out[3] = average duration of one
 * launch of the dominant traversal kernel (persistent strategy: a persistent
 * round; level-synchronous: a breadth-first level).  out[4..12): with
 * KDOP_OPT_KERNEL_TIMING 2, the time of each persistent round (zeros past the
 * last round).  All zeros with KDOP_OPT_KERNEL_TIMING 0 (the default).  out
 * holds 12 doubles. */
int kdop_last_timing(kdop_ctx* ctx, double* out);

/* Gather roof of this context's node pool (measurement; no counterpart in
 * the reference): nloads random records (uniform over the first `window`
 * node slots, 0 = the whole pool; seeded by `seed`) loaded the way the
 * traversal's 4-lane expansion loads them, 4 per quad at a time, with every
 * resident wave of the GPU.  *gbytes_per_s receives the median rate over six
 * runs, counting each record's used words (16 bytes x ceil((K + 2) / 4)).  A
 * small window keeps the records in the L2 caches, the whole pool of a large
 * body in the 256 MB last-level cache.  Allocates and frees its index list. */
int kdop_gather_roof(kdop_ctx* ctx, uint64_t window, uint32_t nloads, uint32_t seed, double* gbytes_per_s);

/* Scheduling statistics of the last traversal (persistent strategy; zeros for
 * the level-synchronous one), 20 entries: out[0] breadth-first levels before
 * the persistent rounds, out[1] frontier handed to them, out[2] rounds run,
 * out[3] loop iterations (sum over waves, or over workgroups where waves
 * share a stack), out[4] items expanded, out[5] most iterations of one wave
 * or workgroup in one round, out[6] frontier chunks taken, out[7] persistent
 * waves, out[8] node-pair tests made by the persistent rounds, out[9]
 * persistent launches, out[12] items the tail hand-off passed on, out[15]
 * breadth-first levels launched after the persistent rounds ran out (the
 * spill finished level-synchronously), out[16] node records the rounds'
 * expansions loaded (one per record and item), out[17] distinct records the
 * rounds read (KDOP_OPT_COUNT_RECORDS calls only, else 0); out[10] LDS stack
 * entries per wave, out[11] waves per CU, out[13] lanes per work item and
 * out[14] waves per workgroup in effect (set for either strategy); out[18],
 * out[19] 0. */
int kdop_traversal_stats(kdop_ctx* ctx, uint64_t* out);

/* Options of a live context (kdop_set_option / kdop_get_option).  The
 * traversal knobs take effect at the next kdop_traverse; 0 restores the
 * library default of a knob. */
#define KDOP_OPT_TRAVERSAL 1      /* KDOP_TRAVERSAL_*                                        */
#define KDOP_OPT_QUEUE_DEPTH 2    /* LDS stack entries per wave: 256, 512, 1024 or 2048 (the    */
                                  /*   quad kernel: 128, 256 or 512); 0: the kernel's default,   */
                                  /*   128 (quad) or 512                                        */
#define KDOP_OPT_WAVES_PER_CU 3   /* persistent waves per CU, 1..32; 0: 28 (quad) or 8          */
#define KDOP_OPT_RAMP_TARGET 4    /* frontier width that ends the breadth-first levels (0:    */
                                  /*   16 items per persistent wave, 64 for calls of 256 pairs */
                                  /*   or more)                                                  */
#define KDOP_OPT_STACK_KEEP 5     /* persistent rounds: once the frontier is gone, a workgroup  */
                                  /*   whose stack holds more than this many entries per wave   */
                                  /*   plus one hand-off hands its bottom 64 entries per wave  */
                                  /*   to the next round; 0 (the default): never                */
#define KDOP_OPT_SEED_BUDGET 6    /* chunk-root pairs tested to seed one call (0: 65,536, or  */
                                  /*   262,144 for calls of 256 pairs or more)                 */
#define KDOP_OPT_KERNEL_TIMING 7  /* HIP events for kdop_last_timing: 0 none (default; the     */
                                  /*   timings read 0), 1 phase events, 2 also one pair on each */
                                  /*   persistent round's dispatch; each event costs a few us   */
#define KDOP_OPT_BFS_MAX 8        /* most breadth-first levels before the persistent rounds; */
                                  /*   0 (the default): by the call's size (1 below 32k tets  */
                                  /*   over the call's pairs, 4 below 400k, else 5)           */
#define KDOP_OPT_NARROW_PATH 10   /* narrow phase: 0 auto (by the last surface-pair count), */
                                  /*   1 one wave per pair, 2 SAT gate (8 lanes per pair) then */
                                  /*   one wave per surviving pair; same results either way     */
#define KDOP_OPT_GRAPHS 11        /* 1: kdop_traverse replays its launches from a HIP graph    */
                                  /*   recorded per set of launch arguments; 0 (default):      */
                                  /*   direct launches                                         */
#define KDOP_OPT_ASYNC_REFIT 9    /* 1: kdop_refit / kdop_refit_device return once the work is */
                                  /*   queued on the context stream (later calls order after it; */
                                  /*   device pose buffers must stay valid until the next call   */
                                  /*   that synchronises, e.g. kdop_traverse)                    */
#define KDOP_OPT_SINGLE_ITEMS 12  /* most (sphere, node) items in one breadth-first step of     */
                                  /*   kdop_single_traversal before it splits the spheres into   */
                                  /*   batches (0: the default, 2^28; 1 .. 2^30); bounds its     */
                                  /*   device memory at 8 bytes x 2 x this, results unchanged    */
#define KDOP_OPT_BLOCK_WAVES 13   /* persistent rounds: waves per workgroup sharing one LDS      */
                                  /*   stack.  Quad kernel: 2, 4 (0: the default) or 8, stack of  */
                                  /*   KDOP_OPT_QUEUE_DEPTH entries per wave.  One lane per item: */
                                  /*   2 (0: the default), 4 or 8 with 512 entries per wave, or 1 */
                                  /*   (one stack per wave of KDOP_OPT_QUEUE_DEPTH entries).      */
                                  /*   Results unchanged                                          */
#define KDOP_OPT_ITEM_LANES 14    /* persistent rounds: lanes per work item.  1: one lane loads    */
                                  /*   and tests whole records; 4: the 4 lanes of a quad each load */
                                  /*   a quarter of every record (coalesced 64-byte runs) and own   */
                                  /*   one child combination; the stack then holds                 */
                                  /*   KDOP_OPT_QUEUE_DEPTH (256 or 512) items per wave.  With 4   */
                                  /*   and K >= 12, the breadth-first levels of both strategies   */
                                  /*   and the persistent strategy's seeds take 4 lanes per item   */
                                  /*   too.  Results unchanged                                     */
#define KDOP_OPT_ROUND_ITERS 18   /* quad kernel: once the frontier is gone, a workgroup that   */
                                  /*   has run this many iterations in a round hands its whole  */
                                  /*   stack to the next round (spread there over every         */
                                  /*   workgroup) and stops; by default 24 for calls of 64 pairs */
                                  /*   or more and 64 below (kdop_get_option reports the last    */
                                  /*   call's), 0 means no budget.  The last round allowed       */
                                  /*   (KDOP_OPT_MAX_ROUNDS)                                      */
                                  /*   runs unbudgeted.  Results unchanged                      */
#define KDOP_OPT_QUAD_ITEMS 17    /* quad kernel: items each quad expands per iteration, 1 (0:  */
                                  /*   the default) or 2 (twice as wide an iteration, 4-wave     */
                                  /*   workgroups).  Results unchanged                           */
#define KDOP_OPT_NODE_SLAB 19     /* quad kernel: 1 stages each wave's node records in an LDS    */
                                  /*   node slab with whole-line loads (8 lanes per record) and   */
                                  /*   tests from LDS; fewer workgroups per CU fit.  0 (default): */
                                  /*   the quads load their words straight into registers.        */
                                  /*   Results unchanged                                           */
#define KDOP_OPT_COUNT_RECORDS 16 /* 1: measurement calls -- the persistent rounds (quad kernel,  */
                                  /*   default shape) also mark each node record they read, and   */
                                  /*   kdop_traversal_stats reports how many distinct ones; slower */
                                  /*   (not for timing), results unchanged.  0: off (default)     */
#define KDOP_OPT_MAX_ROUNDS 15    /* persistent strategy: most persistent rounds per call, 1..7   */
                                  /*   (0: 7); a spill left after them becomes a breadth-first     */
                                  /*   level and the level-synchronous launches finish the        */
                                  /*   traversal from it, however deep.  Results unchanged        */

/* Set / read one option (no counterpart in the reference, whose OpenCL
 * path fixes these sizes at compile time).  Returns KDOP_ERR_ARG for an
 * unknown option or a value out of range. */
int kdop_set_option(kdop_ctx* ctx, int option, int64_t value);
int kdop_get_option(kdop_ctx* ctx, int option, int64_t* value);

/* Autotuner: picks the traversal options (strategy, lanes per work item,
 * round iteration budget, waves per persistent workgroup, waves per CU, LDS
 * queue depth, stack keep, breadth-first target, seed budget) for this context by timing
 * kdop_traverse(npairs, body_pairs, flags) itself: coordinate descent over
 * the options in that order, each candidate measured as the median host wall
 * time of 5 calls and taken only when it beats the best so far by more than
 * 1 % (a candidate that fails with KDOP_ERR_CAPACITY is skipped),
 * until budget_ms of tuning time is spent.  The best options stay set on the
 * context, and the last traversal's results are those of a run with them.
 * chosen (may be NULL) receives KDOP_AUTOTUNE_CHOSEN values indexed by
 * KDOP_OPT_* - 1 (the value in effect of every option; 0 for numbers that
 * name no option); *best_ms (may be NULL) the median time of the best.
 * No counterpart in the reference (its OpenCL sizes are compile-time). */
#define KDOP_AUTOTUNE_CHOSEN 20
int kdop_autotune(kdop_ctx* ctx, uint32_t npairs, const uint32_t* body_pairs, uint32_t flags, double budget_ms, int64_t* chosen,
                  double* best_ms);

/* Stream of the context (hipStream_t), for callers that order their own work. */
void* kdop_stream(kdop_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* KDOP_AMD_H */
