// Discretization graph on the device: node adjacency (the clique union of the
// elements, columns sorted), node -> element incidences, and the element-pair
// position tables the assembly kernels address the CSR values with.  The same
// arrays as goal::build_graph (goal_mesh.cpp), built from radix sorts:
//   - incidences: node ids of every (element, local node) slot, stably sorted
//     with the slot index as value, so each node's slots stay in ascending
//     element order;
//   - adjacency: one 64-bit key (a << nb | b, nb the bits of a node id) per
//     ordered local node pair of every element, sorted and made unique; node
//     a's row is the run of keys with high part a, and its columns come out
//     ascending;
//   - pos[(e npe + a) npe + b] = position of conn[e][b] in node conn[e][a]'s
//     row, rec[inc[k]] = k, posi[k][b] = pos[inc[k]][b].
// Row starts are binary searches in the sorted keys, so nothing depends on
// the order of atomics: the arrays equal the host build's bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "goal_kernels.hpp"

namespace goal {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

__global__ __launch_bounds__(kBlock) void pair_keys_kernel(const int32_t* conn, int64_t ne, int npe, int nb,
                                                           uint64_t* keys) {
  const int64_t n = ne * npe * npe;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t e = i / (npe * npe);
    const int r = (int)(i % (npe * npe)), a = r / npe, b = r % npe;
    keys[i] = ((uint64_t)(uint32_t)conn[e * npe + a] << nb) | (uint32_t)conn[e * npe + b];
  }
}

__global__ __launch_bounds__(kBlock) void iota_kernel(int32_t* v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    v[i] = (int32_t)i;
}

// ptr[a] = first i with key(i) >= a, for a = 0..nn (keys sorted ascending)
template <typename K, typename F>
__device__ inline int64_t lower_bound_by(const K* keys, int64_t n, int64_t a, F key_of) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (key_of(keys[mid]) < a) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(kBlock) void row_starts_u64_kernel(const uint64_t* keys, int64_t n, int64_t nn, int nb,
                                                                int64_t* ptr) {
  for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a <= nn; a += (int64_t)gridDim.x * kBlock)
    ptr[a] = lower_bound_by(keys, n, a, [nb](uint64_t k) { return (int64_t)(k >> nb); });
}

__global__ __launch_bounds__(kBlock) void row_starts_i32_kernel(const int32_t* keys, int64_t n, int64_t nn,
                                                                int64_t* ptr) {
  for (int64_t a = (int64_t)blockIdx.x * kBlock + threadIdx.x; a <= nn; a += (int64_t)gridDim.x * kBlock)
    ptr[a] = lower_bound_by(keys, n, a, [](int32_t k) { return (int64_t)k; });
}

__global__ __launch_bounds__(kBlock) void low_part_kernel(const uint64_t* keys, int64_t n, int nb, int32_t* adj) {
  const uint64_t mask = (uint64_t(1) << nb) - 1;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    adj[i] = (int32_t)(keys[i] & mask);
}

__global__ __launch_bounds__(kBlock) void pos_kernel(const int32_t* conn, int64_t ne, int npe, const int64_t* ptr,
                                                     const int32_t* adj, uint16_t* pos) {
  const int64_t n = ne * npe * npe;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t e = i / (npe * npe);
    const int r = (int)(i % (npe * npe)), a = r / npe, b = r % npe;
    const int32_t na = conn[e * npe + a], nb = conn[e * npe + b];
    const int64_t lo = ptr[na], hi = ptr[na + 1];
    int64_t l = lo, h = hi;
    while (l < h) {
      const int64_t mid = (l + h) >> 1;
      if (adj[mid] < nb) l = mid + 1;
      else h = mid;
    }
    pos[i] = (uint16_t)(l - lo);
  }
}

__global__ __launch_bounds__(kBlock) void rec_posi_kernel(const int32_t* inc, int64_t ninc, int npe,
                                                          const uint16_t* pos, int32_t* rec, uint16_t* posi) {
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < ninc; k += (int64_t)gridDim.x * kBlock) {
    const int32_t s = inc[k];
    rec[s] = (int32_t)k;
    for (int b = 0; b < npe; ++b) posi[k * npe + b] = pos[(int64_t)s * npe + b];
  }
}

int bits_for(int64_t n) {  // bits needed for the values 0..n
  int b = 1;
  while (b < 31 && (int64_t(1) << b) <= n) ++b;
  return b;
}

}  // namespace

hipError_t graph_temp_bytes(int64_t ne, int npe, int64_t nn, size_t* bytes) {
  const int64_t np = ne * npe * npe, ns = ne * npe;
  size_t a = 0, b = 0, c = 0;
  hipError_t err;
  err = hipcub::DeviceRadixSort::SortKeys(nullptr, a, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)np, 0,
                                          2 * bits_for(nn));
  if (err != hipSuccess) return err;
  err = hipcub::DeviceSelect::Unique(nullptr, b, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int64_t*)nullptr,
                                     (int)np);
  if (err != hipSuccess) return err;
  err = hipcub::DeviceRadixSort::SortPairs(nullptr, c, (const int32_t*)nullptr, (int32_t*)nullptr,
                                           (const int32_t*)nullptr, (int32_t*)nullptr, (int)ns, 0, bits_for(nn));
  if (err != hipSuccess) return err;
  *bytes = std::max(a, std::max(b, c));
  return hipSuccess;
}

hipError_t graph_adjacency(const int32_t* conn, int64_t ne, int npe, int64_t nn, void* temp, size_t temp_bytes,
                           uint64_t* keys, uint64_t* keys_sorted, int64_t* num_unique, hipStream_t s) {
  const int64_t np = ne * npe * npe;
  const int nb = bits_for(nn);
  hipLaunchKernelGGL(pair_keys_kernel, dim3(grid_for(np)), dim3(kBlock), 0, s, conn, ne, npe, nb, keys);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  size_t tb = temp_bytes;
  err = hipcub::DeviceRadixSort::SortKeys(temp, tb, keys, keys_sorted, (int)np, 0, 2 * nb, s);
  if (err != hipSuccess) return err;
  tb = temp_bytes;
  return hipcub::DeviceSelect::Unique(temp, tb, keys_sorted, keys, num_unique, (int)np, s);
}

hipError_t graph_rows(const uint64_t* ukeys, int64_t nu, int64_t nn, int64_t* adj_ptr, int32_t* adj, hipStream_t s) {
  const int nb = bits_for(nn);
  hipLaunchKernelGGL(row_starts_u64_kernel, dim3(grid_for(nn + 1)), dim3(kBlock), 0, s, ukeys, nu, nn, nb, adj_ptr);
  hipLaunchKernelGGL(low_part_kernel, dim3(grid_for(nu)), dim3(kBlock), 0, s, ukeys, nu, nb, adj);
  return hipGetLastError();
}

hipError_t graph_incidence(const int32_t* conn, int64_t ne, int npe, int64_t nn, void* temp, size_t temp_bytes,
                           int32_t* keys_out, int32_t* slots, int32_t* inc, int64_t* inc_ptr, hipStream_t s) {
  const int64_t ns = ne * npe;
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(ns)), dim3(kBlock), 0, s, slots, ns);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  size_t tb = temp_bytes;
  err = hipcub::DeviceRadixSort::SortPairs(temp, tb, conn, keys_out, slots, inc, (int)ns, 0, bits_for(nn), s);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(row_starts_i32_kernel, dim3(grid_for(nn + 1)), dim3(kBlock), 0, s, keys_out, ns, nn, inc_ptr);
  return hipGetLastError();
}

hipError_t graph_positions(const int32_t* conn, int64_t ne, int npe, const int64_t* adj_ptr, const int32_t* adj,
                           const int32_t* inc, uint16_t* pos, int32_t* rec, uint16_t* posi, hipStream_t s) {
  const int64_t np = ne * npe * npe, ns = ne * npe;
  hipLaunchKernelGGL(pos_kernel, dim3(grid_for(np)), dim3(kBlock), 0, s, conn, ne, npe, adj_ptr, adj, pos);
  hipLaunchKernelGGL(rec_posi_kernel, dim3(grid_for(ns)), dim3(kBlock), 0, s, inc, ns, npe, pos, rec, posi);
  return hipGetLastError();
}

}  // namespace goal
