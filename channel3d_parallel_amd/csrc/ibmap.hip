// Device construction of the spreading maps of the immersed-boundary
// operator R (src/immersed_boundary_operators.f90:544-603) for bodies whose
// markers move every Runge-Kutta stage.  The map lists, for every field node
// that some marker's 5x5x5 stencil touches, the stencil entries that land
// on it in ascending (marker, point) order, so that each node's sum is
// formed in the reference's order; the nodes are ranked by descending entry
// count (jagged diagonal layout, see Spread in ib.hpp).
//
// Steps, all on the stream, with no host round trip:
//   1. key = node offset, value = entry e = l*125 + p, for every entry;
//   2. stable radix sort by key (entries of one node stay in ascending e);
//   3. run-length encoding: the distinct nodes, ascending, and their counts;
//   4. exclusive scan of the counts: where each node's entries start;
//   5. stable radix sort of the counts, descending, carrying the node's run
//      number (ties stay in ascending offset, as std::stable_sort does on the
//      host);
//   6. per[i] = number of ranks with more than i entries, its exclusive scan
//      gives slot[i], the start of the i-th entries of all ranks;
//   7. scatter: entry i of rank r goes to slot[i] + r.
// Unused tail slots of the count arrays are zero, so every pass runs over
// the fixed capacity nb*125 and the node count stays on the device.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "ib.hpp"

namespace c3d {
namespace ib {
namespace {

constexpr int kMapBlock = 256;

inline dim3 map_blocks(size_t n) { return dim3((unsigned)((n + kMapBlock - 1) / kMapBlock)); }

// 1: node offsets of the entries of component comp (stencil comp: 0 u, 1 v, 2 w).
__global__ void k_map_keys(Stencil st, int nb, int n1, int n2, int* __restrict__ key,
                           int* __restrict__ val) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nb * 125) return;
  const int l = e / 125, p = e - 125 * l;
  const int i = p % 5, j = (p / 5) % 5, k = p / 25;
  const int X = st.idx[(0 * 5 + i) * nb + l], Y = st.idx[(1 * 5 + j) * nb + l],
            Z = st.idx[(2 * 5 + k) * nb + l];
  key[e] = (int)fidx(X, Y, Z, n1, n2);
  val[e] = e;
}

__global__ void k_map_fill(int* __restrict__ a, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_map_iota(int* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

// 6: ranks r cover the counts jcnt[r+1] <= i < jcnt[r] (jcnt non-increasing,
// zero past the last node).
__global__ void k_map_per(const int* __restrict__ jcnt, int n, int* __restrict__ per) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int c = jcnt[r], c1 = (r + 1 < n) ? jcnt[r + 1] : 0;
  for (int i = c1; i < c; i++) per[i] = r + 1;
}

// 7: rank r <- run order[r]: node offset, count and the entries.
__global__ void k_map_scatter(const int* __restrict__ jcnt, const int* __restrict__ order,
                              const int* __restrict__ uniq, const int* __restrict__ start,
                              const int* __restrict__ sorted_e, const int* __restrict__ slot, int n,
                              int* __restrict__ node, int* __restrict__ cnt, int* __restrict__ q) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int c = jcnt[r];
  if (c == 0) return;
  const int src = order[r];
  node[r] = uniq[src];
  cnt[r] = c;
  const int b = start[src];
  for (int i = 0; i < c; i++) q[slot[i] + r] = sorted_e[b + i];
}

// Scratch layout: 11 int arrays of n = nb*125 (+1) entries, then the
// library's temporary storage.
struct Work {
  int *key, *key_s, *val, *val_s, *uniq, *counts, *start, *iota, *jcnt, *order, *per;
  void* temp;
  size_t temp_bytes;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Radix passes cover only the bits a key can hold: node offsets below nf,
// node counts at most 4*nb (a marker lands on one node at most twice per
// periodic axis, when its 5-point window wraps a period of 4 in x and z).
int bits_for(size_t v) {
  int b = 1;
  while (b < 31 && ((size_t)1 << b) <= v) b++;
  return b;
}

size_t cub_temp_bytes(int n, int key_bits, int count_bits) {
  size_t need = 0, b = 0;
  int *i0 = nullptr, *i1 = nullptr, *i2 = nullptr, *i3 = nullptr, *i4 = nullptr;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, i0, i1, i2, i3, n, 0, key_bits);
  need = std::max(need, b);
  (void)hipcub::DeviceRunLengthEncode::Encode(nullptr, b, i0, i1, i2, i3, n);
  need = std::max(need, b);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, i0, i1, n + 1);
  need = std::max(need, b);
  (void)hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, i0, i1, i2, i4, n, 0, count_bits);
  need = std::max(need, b);
  return need;
}

Work carve(void* base, int nb, size_t nf) {
  const size_t n = (size_t)nb * 125 + 1, stride = align256(n * sizeof(int));
  char* p = static_cast<char*>(base);
  Work w{};
  int** arr[11] = {&w.key, &w.key_s, &w.val, &w.val_s, &w.uniq, &w.counts,
                   &w.start, &w.iota, &w.jcnt, &w.order, &w.per};
  for (int a = 0; a < 11; a++) *arr[a] = reinterpret_cast<int*>(p + a * stride);
  w.temp = p + 11 * stride;
  w.temp_bytes = cub_temp_bytes(nb * 125, bits_for(nf), bits_for(4 * (size_t)nb));
  return w;
}

}  // namespace

size_t spread_map_work_bytes(int nb, size_t nf) {
  const size_t n = (size_t)nb * 125 + 1;
  size_t temp = 0;  // the largest of the three components' needs (nf: the largest field)
  for (size_t f = 1; f <= nf; f <<= 1) {
    temp = std::max(temp, cub_temp_bytes(nb * 125, bits_for(f), bits_for(4 * (size_t)nb)));
    if (bits_for(f) >= bits_for(nf)) break;
  }
  temp = std::max(temp, cub_temp_bytes(nb * 125, bits_for(nf), bits_for(4 * (size_t)nb)));
  return 11 * align256(n * sizeof(int)) + align256(temp);
}

hipError_t build_spread_map(const Markers& m, int comp, int n1, int n2, size_t nf, void* work,
                            Spread& sp, int* nnodes, hipStream_t s) {
  const int nb = m.nb, n = nb * 125;
  Work w = carve(work, nb, nf);
  const int kb = bits_for(nf), cb = bits_for(4 * (size_t)nb);
  hipError_t e;
  hipLaunchKernelGGL(k_map_keys, map_blocks(n), dim3(kMapBlock), 0, s, m.st[comp], nb, n1, n2,
                     w.key, w.val);
  size_t tb = w.temp_bytes;
  if ((e = hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.key, w.key_s, w.val, w.val_s, n, 0,
                                               kb, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_map_fill, map_blocks(n + 1), dim3(kMapBlock), 0, s, w.counts, 0, n + 1);
  tb = w.temp_bytes;
  if ((e = hipcub::DeviceRunLengthEncode::Encode(w.temp, tb, w.key_s, w.uniq, w.counts, nnodes, n,
                                                  s)) != hipSuccess)
    return e;
  tb = w.temp_bytes;
  if ((e = hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.counts, w.start, n, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_map_iota, map_blocks(n), dim3(kMapBlock), 0, s, w.iota, n);
  tb = w.temp_bytes;
  if ((e = hipcub::DeviceRadixSort::SortPairsDescending(w.temp, tb, w.counts, w.jcnt, w.iota,
                                                         w.order, n, 0, cb, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_map_fill, map_blocks(n + 1), dim3(kMapBlock), 0, s, w.per, 0, n + 1);
  hipLaunchKernelGGL(k_map_per, map_blocks(n), dim3(kMapBlock), 0, s, w.jcnt, n, w.per);
  tb = w.temp_bytes;
  int* slot = const_cast<int*>(sp.slot);
  if ((e = hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.per, slot, n + 1, s)) != hipSuccess)
    return e;
  hipLaunchKernelGGL(k_map_scatter, map_blocks(n), dim3(kMapBlock), 0, s, w.jcnt, w.order, w.uniq,
                     w.start, w.val_s, slot, n, const_cast<int*>(sp.node),
                     const_cast<int*>(sp.cnt), const_cast<int*>(sp.q));
  return hipGetLastError();
}

}  // namespace ib
}  // namespace c3d
