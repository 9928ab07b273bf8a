/*
 * kdop_meshgen.h -- host-side mesh preparation for the kDOP library (C ABI).
 *
 * These run on the CPU only (no device needed).  They produce the inputs the
 * reference prepares before make_tree: a Morton-ordered tetrahedral mesh and
 * its surface map.  The box-grid tetrahedraliser and the seeded random stream
 * are this library's own synthetic-input generators (the reference authors
 * meshes with TetGen, SIMULATION/PROCEDURAL, which is not part of this path).
 */
#ifndef KDOP_MESHGEN_H
#define KDOP_MESHGEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vertex and tetrahedron counts of an n x n x n box grid: (n+1)^3 and 5 n^3. */
void kdop_box_mesh_size(uint32_t n, uint32_t* nv, uint32_t* nt);

/* Tetrahedralise the box [-size/2, size/2]^3 into n^3 cubes of 5 positively
 * oriented tetrahedra each (cube parity alternates, so faces conform).
 * Interior vertices get N(0, (jitter*h)^2) offsets per axis (h = size/n) from
 * the seeded stream; a draw that would shrink an incident tetrahedron below 10%
 * of its undeformed volume is redrawn.  Output is in grid order (not reordered).
 * X, Y, Z: (n+1)^3 floats; tets: 4 * 5n^3 uint32.  Returns 0 or -1. */
int kdop_make_box_mesh(uint32_t n, float size, double jitter, uint64_t seed, float* X, float* Y, float* Z,
                       uint32_t* tets);

/* Morton reorder of tetrahedra by centroid, then vertices renumbered in first-
 * touch order.  Restates kdop::mesh_reorder (SIMULATION/KDOP/KDOP/include/
 * kdop_mesh_reorder.h:172-353) with one change: equal Morton codes keep their
 * input order (stable sort), where the reference's std::sort leaves ties in an
 * implementation-defined order.  Outputs must not alias inputs. */
int kdop_mesh_reorder(uint32_t nv, const float* X, const float* Y, const float* Z, uint32_t nt, const uint32_t* tets,
                      float* Xo, float* Yo, float* Zo, uint32_t* tets_o);

/* Surface map: surface[t] bit v set when the face of tet t opposite its v-th
 * vertex has no neighbouring tetrahedron.  Restates mesh_array::compute_surface_map
 * (FOUNDATION/MESH_ARRAY/MESH_ARRAY/include/mesh_array_compute_surface_map.h:41-76). */
int kdop_compute_surface_map(uint32_t nv, uint32_t nt, const uint32_t* tets, uint8_t* surface);

/* Chunked-tree shape for (K, chunk_bytes, nt) by the reference formula
 * (kdop_make_tree.h:216-229): out = {node_bytes, N_perfect, L, C, H}.  Returns
 * -1 when the chunk budget cannot hold a two-leaf chunk. */
int kdop_tree_shape(int32_t K, uint64_t chunk_bytes, uint64_t nt, uint64_t* out);

/* Seeded random stream (SplitMix64-seeded xoshiro256**): n doubles uniform in
 * [0,1) from (seed, stream). */
void kdop_rng_uniform(uint64_t seed, uint64_t stream, uint64_t n, double* out);
/* n standard normal doubles (Box-Muller on the same stream). */
void kdop_rng_normal(uint64_t seed, uint64_t stream, uint64_t n, double* out);
/* n uniformly distributed unit quaternions (Shoemake), 4 floats each (x, y, z, w). */
void kdop_rng_quaternions(uint64_t seed, uint64_t stream, uint64_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* KDOP_MESHGEN_H */
