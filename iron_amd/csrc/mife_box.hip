// Device structures of a generated box: node coordinates, element connectivity,
// the equations CSR, the Dirichlet reduction and the SELL-64 storage of the
// reduced (free x free) matrix, each written from closed forms of the grid
// indices.  No sort, no host round trip and no allocation: the sizes follow on
// the host from the box alone (mife_box_sizes), so a caller can allocate,
// launch and go on while the device builds.
//
// Replaces, for a box, iron's generated regular mesh and its DOF numbering
// (generated_mesh_routines.F90:1568-1612, :3217-3384; field_routines.F90:10056),
// the equations-matrix structure (EquationsMatrix_StructureCalculate,
// equations_matrices_routines.F90:4316-4602) and the solver mapping and solver
// matrix structure (solver_mapping_routines.F90:422-547, :1445-1511;
// solver_matrices_routines.F90:1612-2431): solver rows and columns are the free
// DOFs in ascending global order.
//
// Numbering (as mesh.py and the generic kernels number a generated mesh):
// node (ix, iy, iz) = ix + Nx (iy + Ny iz), elements likewise, local node
// a = ax + NB (ay + NB az); DOF (c, node) = c N + node (separated order); the
// CSR row of a DOF lists every DOF of every node it shares an element with,
// ascending: component c2, then z, y, x neighbour.  Dirichlet planes: x = 0
// (fix_lo) and x = x_max (fix_hi), every component.
#include "mife_common.hpp"
#include "mife_grid.hpp"

#include <algorithm>

namespace mife {
namespace {

struct Box {
  int P, NC, nx, ny, nz, Nx, Ny, Nz, N;
  int lo, hi, nfx;  // fixed x-planes (0/1) and free nodes per x-line
  int Sx, Sy, Sz;   // coupled (node, neighbour) pairs per x-, y- and z-line
  int W;            // SELL width: the longest reduced row (SELL3: in node blocks)
  int block;        // SELL3: free DOFs node by node, 3x3 node blocks
  long long n_free, nnz;
};

}  // namespace

// Whether node plane ix is a Dirichlet plane.
__host__ __device__ inline bool box_x_fixed(int lo, int hi, int Nx, int ix) {
  return (lo && ix == 0) || (hi && ix == Nx - 1);
}

static int make_box(const mife_box_desc* d, Box* B, mife_box_counts* out) {
  MIFE_REQUIRE(d, MIFE_ERR_POINTER_IS_NULL, "null box description");
  MIFE_REQUIRE(d->order == 1 || d->order == 2, MIFE_ERR_ARG, "box order %d (1 trilinear, 2 triquadratic)", d->order);
  MIFE_REQUIRE(d->n_comp == 1 || d->n_comp == 3, MIFE_ERR_ARG, "%d components (1 or 3)", d->n_comp);
  MIFE_REQUIRE(d->n_elem[0] >= 1 && d->n_elem[1] >= 1 && d->n_elem[2] >= 1, MIFE_ERR_ARG, "empty box");
  MIFE_REQUIRE((d->fix_lo == 0 || d->fix_lo == 1) && (d->fix_hi == 0 || d->fix_hi == 1), MIFE_ERR_ARG,
               "fix_lo / fix_hi are 0 or 1");
  Box b{};
  b.P = d->order;
  b.NC = d->n_comp;
  b.nx = d->n_elem[0];
  b.ny = d->n_elem[1];
  b.nz = d->n_elem[2];
  const long long Nx = (long long)b.P * b.nx + 1, Ny = (long long)b.P * b.ny + 1, Nz = (long long)b.P * b.nz + 1;
  MIFE_REQUIRE(b.NC * Nx * Ny * Nz < (1LL << 31), MIFE_ERR_ARG, "box of %lld DOFs exceeds the int32 range",
               b.NC * Nx * Ny * Nz);
  b.Nx = (int)Nx;
  b.Ny = (int)Ny;
  b.Nz = (int)Nz;
  b.N = (int)(Nx * Ny * Nz);
  b.lo = d->fix_lo;
  b.hi = d->fix_hi;
  b.nfx = b.Nx - b.lo - b.hi;
  MIFE_REQUIRE(b.nfx >= 1, MIFE_ERR_ARG, "every node of the box is fixed");
  b.Sx = x_pair_start(b.Nx, b.P, b.nx);
  b.Sy = x_pair_start(b.Ny, b.P, b.ny);
  b.Sz = x_pair_start(b.Nz, b.P, b.nz);
  // reduced pairs along x (free node, free neighbour) and the longest reduced row
  long long sxf = 0;
  int mx = 0, my = 0, mz = 0;
  for (int ix = 0; ix < b.Nx; ++ix) {
    if (box_x_fixed(b.lo, b.hi, b.Nx, ix)) continue;
    int lx, hx, f = 0;
    elem_range(ix, b.P, b.nx, lx, hx);
    for (int jx = b.P * lx; jx <= b.P * hx + b.P; ++jx) f += box_x_fixed(b.lo, b.hi, b.Nx, jx) ? 0 : 1;
    sxf += f;
    mx = f > mx ? f : mx;
  }
  for (int i = 0; i < b.Ny; ++i) my = neighbour_count(i, b.P, b.ny) > my ? neighbour_count(i, b.P, b.ny) : my;
  for (int i = 0; i < b.Nz; ++i) mz = neighbour_count(i, b.P, b.nz) > mz ? neighbour_count(i, b.P, b.nz) : mz;
  b.n_free = (long long)b.NC * b.nfx * b.Ny * b.Nz;
  b.nnz = (long long)b.NC * b.NC * b.Sx * (long long)b.Sy * b.Sz;
  // a slice's columns (solver numbering of one component, or node numbering)
  // lie within P (nfx Ny + nfx + 1) of its rows
  const bool fits16 = 63 + 2LL * b.P * ((long long)b.nfx * b.Ny + b.nfx + 1) < 65536;
  b.block = (d->block == 1 && b.NC == 3 && fits16) ? 1 : 0;
  b.W = b.block ? mx * my * mz : b.NC * mx * my * mz;
  const long long rows = b.block ? b.n_free / 3 : b.n_free;  // SELL rows: nodes or DOFs
  const long long n_slices = (rows + kWave - 1) / kWave;
  const long long stored = n_slices * kWave * b.W * (b.block ? 9 : 1);
  MIFE_REQUIRE(b.nnz < (1LL << 31) && stored < (1LL << 31), MIFE_ERR_ARG,
               "box matrix of %lld entries (%lld stored) exceeds the int32 range", b.nnz, stored);
  if (B) *B = b;
  if (out) {
    out->n_nodes = b.N;
    out->n_elem = (long long)b.nx * b.ny * b.nz;
    out->nodes_per_elem = (b.P + 1) * (b.P + 1) * (b.P + 1);
    out->n_dofs = (long long)b.NC * b.N;
    out->nnz = b.nnz;
    out->n_free = b.n_free;
    out->red_nnz = (long long)b.NC * b.NC * sxf * b.Sy * b.Sz;
    out->n_slices = n_slices;
    out->width = b.W;
    out->stored = stored;
    out->col16 = (b.block || (b.NC == 1 && fits16)) ? 1 : 0;
    out->block = b.block;
  }
  return MIFE_OK;
}

namespace {

// Node (ix, iy, iz) of node index `node`.
__device__ __forceinline__ void node_ijk(const Box& B, int node, int& ix, int& iy, int& iz) {
  ix = node % B.Nx;
  iy = (node / B.Nx) % B.Ny;
  iz = node / (B.Nx * B.Ny);
}

// Free node index of node (ix, iy, iz) (x fastest over the free x-planes).
__device__ __forceinline__ int box_free_node(const Box& B, int ix, int iy, int iz) {
  return (iz * B.Ny + iy) * B.nfx + (ix - B.lo);
}

// Solver index of DOF (c, ix, iy, iz), or -1 when fixed: component by component
// (free DOFs ascending), or node by node with SELL3.
__device__ __forceinline__ int box_new_idx(const Box& B, int c, int ix, int iy, int iz) {
  if (box_x_fixed(B.lo, B.hi, B.Nx, ix)) return -1;
  if (B.block) return 3 * box_free_node(B, ix, iy, iz) + c;
  return (c * B.Nz + iz) * B.Ny * B.nfx + iy * B.nfx + (ix - B.lo);
}

// First CSR entry of equations row (c, ix, iy, iz): the rows before it, x fastest,
// hold NC c_x c_y c_z entries each (c_d: neighbours along d), a sum that
// separates into the per-direction pair counts.
__device__ __forceinline__ int box_row_start(const Box& B, int c, int ix, int iy, int iz) {
  const int cz = neighbour_count(iz, B.P, B.nz), cy = neighbour_count(iy, B.P, B.ny);
  return B.NC * (c * B.Sx * B.Sy * B.Sz + x_pair_start(iz, B.P, B.nz) * B.Sx * B.Sy +
                 cz * x_pair_start(iy, B.P, B.ny) * B.Sx + cz * cy * x_pair_start(ix, B.P, B.nx));
}

// Per node: coordinates (origin + i * extent / (n - 1), as the generated mesh
// computes them), and per component the solver index, the inverse map and the
// Dirichlet value.
__global__ void box_nodes_kernel(Box B, double hx, double hy, double hz, double u_lo, double u_hi,
                                 double* __restrict__ coords, int32_t* __restrict__ free_idx,
                                 int32_t* __restrict__ new_idx, double* __restrict__ u0) {
#pragma clang fp contract(off)
  for (int node = blockIdx.x * blockDim.x + threadIdx.x; node < B.N; node += gridDim.x * blockDim.x) {
    int ix, iy, iz;
    node_ijk(B, node, ix, iy, iz);
    if (coords) {
      coords[3 * (long long)node + 0] = (double)ix * hx;
      coords[3 * (long long)node + 1] = (double)iy * hy;
      coords[3 * (long long)node + 2] = (double)iz * hz;
    }
    const bool fx = box_x_fixed(B.lo, B.hi, B.Nx, ix);
    for (int c = 0; c < B.NC; ++c) {
      const int row = c * B.N + node;
      const int k = box_new_idx(B, c, ix, iy, iz);
      new_idx[row] = k;
      if (k >= 0) free_idx[k] = row;
      u0[row] = fx ? (B.lo && ix == 0 ? u_lo : u_hi) : 0.0;
    }
  }
}

// Per element: its (P+1)^3 nodes, xi1 fastest.
__global__ void box_elems_kernel(Box B, int32_t* __restrict__ elem_nodes) {
  const int NB = B.P + 1, NN = NB * NB * NB;
  const long long total = (long long)B.nx * B.ny * B.nz * NN;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int e = (int)(t / NN), a = (int)(t - (long long)e * NN);
    const int ex = e % B.nx, ey = (e / B.nx) % B.ny, ez = e / (B.nx * B.ny);
    const int ax = a % NB, ay = (a / NB) % NB, az = a / (NB * NB);
    elem_nodes[t] = (B.P * ex + ax) + B.Nx * ((B.P * ey + ay) + B.Ny * (B.P * ez + az));
  }
}

// One wavefront per equations row: its start and its columns.
__global__ __launch_bounds__(256) void box_csr_kernel(Box B, int32_t* __restrict__ row_ptr,
                                                      int32_t* __restrict__ col) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int n_rows = B.NC * B.N;
  if (row >= n_rows) return;
  const int c = row / B.N;
  int ix, iy, iz;
  node_ijk(B, row - c * B.N, ix, iy, iz);
  const int start = box_row_start(B, c, ix, iy, iz);
  if (lane == 0) {
    row_ptr[row] = start;
    if (row == n_rows - 1) row_ptr[n_rows] = (int)B.nnz;
  }
  int lx, hx, ly, hy, lz, hz;
  elem_range(ix, B.P, B.nx, lx, hx);
  elem_range(iy, B.P, B.ny, ly, hy);
  elem_range(iz, B.P, B.nz, lz, hz);
  const int cx = B.P * (hx - lx) + B.P + 1, cy = B.P * (hy - ly) + B.P + 1, cz = B.P * (hz - lz) + B.P + 1;
  const int cnt = cx * cy * cz;
  for (int p = lane; p < B.NC * cnt; p += 64) {
    const int c2 = p / cnt, q = p - c2 * cnt;
    const int jx = B.P * lx + q % cx, jy = B.P * ly + (q / cx) % cy, jz = B.P * lz + q / (cx * cy);
    col[start + p] = c2 * B.N + jx + B.Nx * (jy + B.Ny * jz);
  }
}

// One wavefront per SELL slice of the reduced matrix, one solver row per lane
// (solver order: free DOFs ascending, no permutation): the row's free columns
// in order, their solver indices and equations-CSR positions, then padding
// (the row's own index, no source) up to the uniform width.  With col16 the
// columns are stored relative to the slice's smallest one.
__global__ __launch_bounds__(256) void box_sell_kernel(Box B, int n_slices, int32_t* __restrict__ slice_ptr,
                                                       int32_t* __restrict__ sell_col, uint16_t* __restrict__ col16,
                                                       int32_t* __restrict__ col_base, int32_t* __restrict__ src) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (s >= n_slices) return;
  const int nf = (int)B.n_free;
  const int k = s * 64 + lane;
  const int base = s * 64 * B.W;
  if (lane == 0) {
    slice_ptr[s] = base;
    if (s == n_slices - 1) slice_ptr[n_slices] = base + 64 * B.W;
  }
  // the row: solver index k -> DOF (c, ix, iy, iz)
  const bool real = k < nf;
  int c = 0, ix = 0, iy = 0, iz = 0;
  if (real) {
    const int per_c = B.nfx * B.Ny * B.Nz;
    c = k / per_c;
    const int r = k - c * per_c;
    iz = r / (B.nfx * B.Ny);
    iy = (r / B.nfx) % B.Ny;
    ix = r % B.nfx + B.lo;
  }
  int lx, hx, ly, hy, lz, hz;
  elem_range(ix, B.P, B.nx, lx, hx);
  elem_range(iy, B.P, B.ny, ly, hy);
  elem_range(iz, B.P, B.nz, lz, hz);
  const int cx = B.P * (hx - lx) + B.P + 1, cy = B.P * (hy - ly) + B.P + 1, cz = B.P * (hz - lz) + B.P + 1;
  // first free column of the row (its smallest): component 0, lowest z, y and
  // free x neighbour
  int jx_first = B.P * lx;
  while (box_x_fixed(B.lo, B.hi, B.Nx, jx_first)) ++jx_first;
  const int first = real ? box_new_idx(B, 0, jx_first, B.P * ly, B.P * lz) : nf - 1;
  int cb = first;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cb = min(cb, __shfl_xor(cb, off, 64));
  if (col16 && lane == 0) col_base[s] = cb;
  auto put = [&](int j, int colv, int srcv) {
    const int e = base + j * 64 + lane;
    if (col16)
      col16[e] = (uint16_t)(colv - cb);
    else
      sell_col[e] = colv;
    src[e] = srcv;
  };
  int j = 0;
  if (real) {
    const int start = box_row_start(B, c, ix, iy, iz);
    int p = 0;
    for (int c2 = 0; c2 < B.NC; ++c2)
      for (int qz = 0; qz < cz; ++qz)
        for (int qy = 0; qy < cy; ++qy)
          for (int qx = 0; qx < cx; ++qx, ++p) {
            const int jx = B.P * lx + qx;
            if (box_x_fixed(B.lo, B.hi, B.Nx, jx)) continue;
            put(j++, box_new_idx(B, c2, jx, B.P * ly + qy, B.P * lz + qz), start + p);
          }
  }
  const int pad = real ? k : nf - 1;
  for (; j < B.W; ++j) put(j, pad, -1);
}

// SELL3 of the reduced matrix (NC = 3, free DOFs node by node): one wavefront
// per slice of 64 free nodes, one node row per lane.  Block j of the row is
// its j-th free neighbour node in (z, y, x) order, stored as a 16-bit offset
// from the slice's smallest neighbour; its nine values come from the
// equations-CSR entries (row (r, node), column (c, neighbour)), t = 3 r + c,
// at src[9 (base + 64 j) + 64 t + lane].  Padding blocks (up to the uniform
// width) point at the row's own node and have no source.
__global__ __launch_bounds__(256) void box_sell3_kernel(Box B, int n_slices, int32_t* __restrict__ slice_ptr,
                                                        uint16_t* __restrict__ col16, int32_t* __restrict__ col_base,
                                                        int32_t* __restrict__ src) {
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (s >= n_slices) return;
  const int nfn = (int)(B.n_free / 3);
  const int k = s * 64 + lane;
  const int base = s * 64 * B.W;  // in blocks
  if (lane == 0) {
    slice_ptr[s] = base;
    if (s == n_slices - 1) slice_ptr[n_slices] = base + 64 * B.W;
  }
  const bool real = k < nfn;
  int ix = 0, iy = 0, iz = 0;
  if (real) {
    ix = k % B.nfx + B.lo;
    iy = (k / B.nfx) % B.Ny;
    iz = k / (B.nfx * B.Ny);
  }
  int lx, hx, ly, hy, lz, hz;
  elem_range(ix, B.P, B.nx, lx, hx);
  elem_range(iy, B.P, B.ny, ly, hy);
  elem_range(iz, B.P, B.nz, lz, hz);
  const int cx = B.P * (hx - lx) + B.P + 1, cy = B.P * (hy - ly) + B.P + 1, cz = B.P * (hz - lz) + B.P + 1;
  int jx_first = B.P * lx;
  while (box_x_fixed(B.lo, B.hi, B.Nx, jx_first)) ++jx_first;
  const int first = real ? box_free_node(B, jx_first, B.P * ly, B.P * lz) : nfn - 1;
  int cb = first;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) cb = min(cb, __shfl_xor(cb, off, 64));
  if (lane == 0) col_base[s] = cb;
  const int per_c = cz * cy * cx;  // equations-CSR entries per column component in the row
  int rs[3] = {0, 0, 0};
  if (real)
#pragma unroll
    for (int r = 0; r < 3; ++r) rs[r] = box_row_start(B, r, ix, iy, iz);
  int j = 0;
  if (real) {
    for (int qz = 0; qz < cz; ++qz)
      for (int qy = 0; qy < cy; ++qy)
        for (int qx = 0; qx < cx; ++qx) {
          const int jx = B.P * lx + qx;
          if (box_x_fixed(B.lo, B.hi, B.Nx, jx)) continue;
          const int b = base + j * 64 + lane;
          col16[b] = (uint16_t)(box_free_node(B, jx, B.P * ly + qy, B.P * lz + qz) - cb);
          const int p = (qz * cy + qy) * cx + qx;
          int32_t* sv = src + (size_t)9 * (base + j * 64) + lane;
#pragma unroll
          for (int t = 0; t < 9; ++t) sv[64 * t] = rs[t / 3] + (t % 3) * per_c + p;
          ++j;
        }
  }
  for (; j < B.W; ++j) {
    col16[base + j * 64 + lane] = (uint16_t)((real ? k : nfn - 1) - cb);
    int32_t* sv = src + (size_t)9 * (base + j * 64) + lane;
#pragma unroll
    for (int t = 0; t < 9; ++t) sv[64 * t] = -1;
  }
}

}  // namespace
}  // namespace mife

using namespace mife;

extern "C" {

int mife_box_sizes(const mife_box_desc* desc, mife_box_counts* out) {
  MIFE_REQUIRE(out, MIFE_ERR_POINTER_IS_NULL, "null output");
  return make_box(desc, nullptr, out);
}

int mife_box_build(const mife_box_desc* desc, const mife_box_arrays* a, void* stream) {
  Box B;
  mife_box_counts n;
  int rc = make_box(desc, &B, &n);
  if (rc) return rc;
  MIFE_REQUIRE(a && a->elem_nodes && a->row_ptr && a->col && a->free_idx && a->new_idx && a->u0 && a->slice_ptr &&
                   a->sell_src,
               MIFE_ERR_POINTER_IS_NULL, "null output array");
  MIFE_REQUIRE(n.col16 ? (a->sell_col16 && a->col_base) : (a->sell_col != nullptr), MIFE_ERR_POINTER_IS_NULL,
               "col16 = %d: %s needed", n.col16, n.col16 ? "sell_col16 and col_base" : "sell_col");
  hipStream_t s = (hipStream_t)stream;
  const double hx = desc->extent[0] / (double)(B.Nx - 1), hy = desc->extent[1] / (double)(B.Ny - 1),
               hz = desc->extent[2] / (double)(B.Nz - 1);
  const int g = std::min(div_up(B.N, 256), 4096);
  box_nodes_kernel<<<g, 256, 0, s>>>(B, hx, hy, hz, desc->u_lo, desc->u_hi, a->coords, a->free_idx, a->new_idx,
                                     a->u0);
  const long long en = n.n_elem * n.nodes_per_elem;
  box_elems_kernel<<<(int)std::min<long long>((en + 255) / 256, 4096), 256, 0, s>>>(B, a->elem_nodes);
  box_csr_kernel<<<div_up(n.n_dofs, 4), 256, 0, s>>>(B, a->row_ptr, a->col);
  if (B.block)
    box_sell3_kernel<<<div_up(n.n_slices, 4), 256, 0, s>>>(B, (int)n.n_slices, a->slice_ptr, a->sell_col16,
                                                           a->col_base, a->sell_src);
  else
    box_sell_kernel<<<div_up(n.n_slices, 4), 256, 0, s>>>(B, (int)n.n_slices, a->slice_ptr,
                                                          n.col16 ? nullptr : a->sell_col,
                                                          n.col16 ? a->sell_col16 : nullptr, a->col_base, a->sell_src);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // extern "C"
