// 1-D neighbour arithmetic of a generated regular grid of order P (nodes and
// elements numbered x fastest, as iron's generated mesh numbers them), shared by
// the structured assembly and the box structure builder (host and device).
#pragma once
#include <hip/hip_runtime.h>

namespace mife {

// 1-D element range of node i (order P, n elements): elements ex with
// P ex <= i <= P ex + P, clipped to [0, n - 1].
__host__ __device__ __forceinline__ void elem_range(int i, int P, int n, int& lo, int& hi) {
  lo = i > 0 ? (i - 1) / P : 0;
  hi = (i / P < n - 1 ? i / P : n - 1);
}

// Element of [lo, hi] (one or two consecutive elements) whose index parity is
// `bit`, or -1.
__host__ __device__ __forceinline__ int elem_of_parity(int lo, int hi, int bit) {
  if ((lo & 1) == bit) return lo;
  if (hi != lo && (hi & 1) == bit) return hi;
  return -1;
}

// Neighbour pairs of an x-line (order P, nx elements): pair t couples node ix to
// node jx, pairs ordered by ix, then jx.  Nodes repeat with period P: a vertex
// node couples to 2P + 1 nodes, each of the P - 1 nodes inside an element to
// P + 1, so a period holds P^2 + 2P pairs; the first vertex lacks its P left
// neighbours, which the shift by P accounts for.  An x-line has
// nx (P^2 + 2P) + 1 pairs.
__host__ __device__ __forceinline__ void x_pair(int t, int P, int& ix, int& jx) {
  const int per = P * P + 2 * P;
  const int tp = t + P;
  const int e = tp / per, r = tp - e * per;
  if (r < 2 * P + 1) {
    ix = P * e;
    jx = ix - P + r;
  } else {
    const int r2 = r - (2 * P + 1), k = r2 / (P + 1);
    ix = P * e + 1 + k;
    jx = P * e + (r2 - k * (P + 1));
  }
}

// First pair of node ix of an x-line with nx elements; ix = P nx + 1 (one past
// the last node, which lacks its P right neighbours) gives the pair count.
__host__ __device__ __forceinline__ int x_pair_start(int ix, int P, int nx) {
  if (ix > P * nx) return nx * (P * P + 2 * P) + 1;
  const int e = ix / P, k = ix - e * P;
  const int tp = e * (P * P + 2 * P) + (k ? 2 * P + 1 + (k - 1) * (P + 1) : 0);
  return tp - P > 0 ? tp - P : 0;
}

// Nodes a node couples to along one direction (order P, n elements).
__host__ __device__ __forceinline__ int neighbour_count(int i, int P, int n) {
  int lo, hi;
  elem_range(i, P, n, lo, hi);
  return P * (hi - lo) + P + 1;
}

}  // namespace mife
