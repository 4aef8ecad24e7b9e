// Device-resident mesh handle shared by the operator entry points and the
// matrix-free solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "assembly.hpp"
#include "common.hpp"
#include "density.hpp"
#include "element.hpp"

struct tatva_mesh {
  int device;
  int kind;
  int dim;
  int nne;
  int64_t n_elements;
  int64_t n_nodes;
  tatva::QuadRule quad;
  double *coords;        // (N, dim) FP64
  int32_t *conn;         // (E, nne) int32, element-major as in tatva
  int32_t *inv_ptr;      // (N + 1) offsets into inv_list
  int32_t *inv_list;     // packed (element << 3 | local node), sorted by element per node
  void *scratch;         // tile partial sums: n_slots * dim doubles
  int64_t scratch_bytes;
  double *partials;      // reduction partials, kMaxReduceBlocks * 4
  int max_incident;      // largest number of elements sharing a node
  int hex_affine;        // Hex8 mesh of parallelepipeds only (constant Jacobian per element)
  // LDS tiling (assembly.hpp)
  tatva::TileSet tiles;
  int64_t n_slots;
  uint16_t *t_lnodes;
  int16_t *t_color_off;
  int32_t *t_node_ptr;
  int32_t *t_slot_ptr;
  int32_t *t_slot_pos;
  int32_t *t_tnodes;
  uint16_t *t_inc_end;
  uint16_t *t_inc_list;
  int32_t *t_elem;
  // point location (2-D, built by the first interpolate): a uniform grid over
  // the nodes' bounding box; cell c lists, ascending, the elements whose
  // bounding box overlaps it in loc_elems[loc_ptr[c] .. loc_ptr[c + 1])
  int32_t *loc_ptr;
  int32_t *loc_elems;
  tatva::PointGrid loc;
};

namespace tatva {

// Element Hessian-vector product assembled into nodes (tile kernel + gather).
// When skip is non-NULL and *skip != 0 at kernel start both kernels return
// immediately (used by graph-captured solver iterations after convergence).
int launch_hvp_f64(tatva_mesh *m, const tatva_density *d, const double *u, const double *v,
                   const int32_t *dof_map, double *y, hipStream_t s, const int *skip = nullptr);
// With lin (a cache written by launch_linearize_f32 at the same u) the tile
// kernel reads the density's linearisation instead of recomputing it.
// With geo (an affine Hex8 mesh's geometry cache, written by
// launch_geometry_cache_f32) J^-1 and det J are read instead of computed.
int launch_hvp_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                   const int32_t *dof_map, float *y, hipStream_t s, const int *skip = nullptr,
                   const float *lin = nullptr, const float *geo = nullptr);
// Floats of the linearisation cache of (mesh, density) at FP32 (0: the density
// has none to cache), and the kernel writing it from the FP32 state u.
int64_t lin_cache_floats(const tatva_mesh *m, const tatva_density *d);
int launch_linearize_f32(tatva_mesh *m, const tatva_density *d, const float *u, float *lin,
                         hipStream_t s, const float *geo = nullptr);
// Floats of an affine Hex8 mesh's geometry cache (0 for other meshes) and its
// writer.
int64_t geo_cache_floats(const tatva_mesh *m);
int launch_geometry_cache_f32(tatva_mesh *m, float *geo, hipStream_t s);
// y = H(u) v (FP32) with the block shares of v.y in part_dot, over
// hvp_gather_blocks(m) blocks (v, y reduced through dof_map); rev: tiles in
// reverse order (same result); between: an event recorded after the tile
// kernel, before the gather (kernel timing); dx: a CG step's deferred p and x
// update, applied by the tile kernel.
int launch_hvp_dot_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                       const int32_t *dof_map, float *y, double *part_dot, hipStream_t s,
                       const int *skip = nullptr, int rev = 0, hipEvent_t between = nullptr,
                       const float *lin = nullptr, const float *geo = nullptr,
                       const CgDefer *dx = nullptr);
int hvp_gather_blocks(const tatva_mesh *m);
int density_ncomp(const tatva_mesh *m, const tatva_density *d);
// Drops cached solver graphs that captured this mesh's arrays.
void forget_mesh_graphs(const tatva_mesh *m);

}  // namespace tatva
