// Voxel-grid passes either side of the SIMP loop: anchor regions, the
// boundary fields they produce, face-connected components, the export iso
// level and the load-path guard (anneal/problem.py:276-326,
// anneal/export.py:8-100).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace anneal {

// Signed-distance primitives (anneal/kernel.py).  Each primitive is one kind
// and 16 doubles of parameters:
//   kRegionBox      c[3] half[3]                      (kernel.py:60-65: axis box)
//   kRegionOBox     o[3] x[3] y[3] z[3] zs half[3]    (kernel.py:66-68: framed box,
//                                                       zs = z0 + size_z / 2)
//   kRegionSphere   c[3] r                            (kernel.py:84-85)
//   kRegionCylinder a[3] ax[3] L r r_in               (kernel.py:106-114; r_in > 0 is a pipe)
enum RegionKind : int { kRegionBox = 0, kRegionOBox = 1, kRegionSphere = 2, kRegionCylinder = 3 };
constexpr int kRegionParams = 16;
constexpr int kMaxAnchors = 32;

// Anchor a selects a point when min over its primitives [first[a], first[a+1])
// of the signed distance is <= thr[a].  Tables live in device memory.
struct RegionTable {
  const int* kind;
  const double* prm;    // [nprim][16]
  const int* first;     // [nanchor + 1]
  const double* thr;    // [nanchor]
  int nanchor;
};

// bits[p] bit a <=> point p is selected by anchor a.  at_nodes: points are the
// grid nodes that belong to a voxel with emask set; otherwise the voxel centres
// with emask set.  counts[a] (device, int64) = number of selected points.
void region_select(const GridDims& g, double h, const double* origin, const RegionTable& t, bool at_nodes,
                   const uint8_t* emask, uint32_t* bits, long long* counts, hipStream_t s);

// out[a][p] = min over anchor a's primitives of the signed distance at pts[p][3].
void region_sdf(int64_t n, const double* pts, const RegionTable& t, double* out, hipStream_t s);

// Per node: fixed-dof bits (OR of dofmask[a] over the anchors that select it)
// and, per load case c, F[c][3 n + d] = sum over anchors in order of load[c][a][d]
// (the evenly spread nodal load v_d / count_a, or 0).
void boundary_fields(const GridDims& g, const uint32_t* node_bits, int nanchor, const uint8_t* dofmask,
                     const double* load, int ncase, uint8_t* fixed_bits, double* F, hipStream_t s);

// ebits[e] = OR of the node bits of voxel e's 8 nodes when emask[e], else 0.
void gather_node_bits(const GridDims& g, const uint32_t* node_bits, const uint8_t* emask, uint32_t* ebits,
                      hipStream_t s);

// Face-connected components of mask: label[e] = smallest voxel index of e's
// component, -1 outside the mask.  Returns the number of components (host
// sync).  counter: 2 device int64 of scratch; the second flags a union step
// that exceeded its bound (reported as an error).
int64_t components(const GridDims& g, const uint8_t* mask, int32_t* label, long long* counter, hipStream_t s);

// Marks the components of mask holding a voxel with ebits & seed; out_mask (if
// not null) = mask restricted to the marked components.  Returns the number of
// marked components; *linked = OR of ebits over the voxels of marked
// components.  keep is nelem bytes of scratch.
int64_t link_components(const GridDims& g, const uint8_t* mask, const int32_t* label, const uint32_t* ebits,
                        uint32_t seed, uint8_t* keep, uint8_t* out_mask, uint32_t* linked, long long* counter,
                        hipStream_t s);

void threshold_mask(int64_t n, const double* f, double t, uint8_t* mask, hipStream_t s);
int64_t count_above(int64_t n, const double* f, double t, long long* counter, hipStream_t s);

// export.output_field (export.py:8-14).
void output_field(int64_t n, const double* x, const double* fill, const uint8_t* active, const double* solid_fill,
                  const double* void_fill, const uint8_t* frozen, double* f, hipStream_t s);

// export.choose_level (export.py:17-28): the 40-step bisection of the iso
// level decided from 4 histogram passes.  work: choose_level_work_bytes().
size_t choose_level_work_bytes();
double choose_level(int64_t n, const double* f, const double* x, const double* fill, void* work, double* target_out,
                    hipStream_t s);

// Shortest paths of the load-path repair (export.py:70-83): voxel graph of the
// active voxels, face neighbours, entering voxel u costs max(0, need - f[u]) +
// 1e-3; sources = active voxels with ebits & source_groups.  dist (FP64) and
// pred (smallest-index predecessor on a shortest path, -1 for sources and
// unreached voxels).  stamp: nelem int32 scratch.  Returns the sweeps used.
int sssp(const GridDims& g, const uint8_t* active, const double* f, double need, const uint32_t* ebits,
         uint32_t source_groups, double* dist, int32_t* pred, int32_t* stamp, int* flag, int max_sweeps,
         hipStream_t s);

// Start voxel (smallest-index minimiser of dist over the active voxels with
// ebits bit `group`) and the path from it back to a source.  Returns the path
// length (0 when unreachable); *start = start voxel (-1 if the group is empty).
int64_t trace_path(const GridDims& g, const uint8_t* active, const uint32_t* ebits, int group, const double* dist,
                   const int32_t* pred, int32_t* path, long long* work, int64_t* start, double* dstart,
                   hipStream_t s);

// export.py:91-94: raise the ring around the path (linear offsets +-1, +-nx,
// +-nx*ny, in range and active) to min(need + 0.05, fill), then the path to
// min(1, fill).
void raise_path(const GridDims& g, const int32_t* path, int64_t len, const uint8_t* active, const double* fill,
                double need, double* f, hipStream_t s);

}  // namespace anneal
