// Device prefix scan and LDS-staged LSD radix sort (32-bit keys, 32-bit
// values), written for wave64: each pass builds per-tile digit histograms in
// LDS, scans them digit-major in HBM, then scatters every tile with stable
// ranks computed from 8 wave ballots per key and per-wave LDS counters.
#include "sort.h"

namespace rtg {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// Exclusive block scan of one value per thread; returns the block total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds_waves[wave] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) {
    uint32_t s = lds_waves[w];
    if (w < wave) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, size_t n,
                                                             uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t lds_waves[SCAN_THREADS / 64];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = (base + k < n) ? in[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, lds_waves, total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                           const uint32_t* __restrict__ tile_offsets) {
  const uint32_t add = tile_offsets[blockIdx.x];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  for (int k = threadIdx.x; k < SCAN_TILE; k += SCAN_THREADS)
    if (base + k < n) out[base + k] += add;
}

void Scanner::exclusive(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  // level sizes
  std::vector<size_t> sizes;
  size_t m = n;
  while (true) {
    size_t tiles = (m + SCAN_TILE - 1) / SCAN_TILE;
    sizes.push_back(tiles);
    if (tiles <= 1) break;
    m = tiles;
  }
  size_t total = 0;
  for (size_t t : sizes) total += 2 * (t + 1);
  tmp.reserve(total * sizeof(uint32_t));
  uint32_t* p = tmp.as<uint32_t>();
  std::vector<uint32_t*> sums, offs;
  for (size_t t : sizes) {
    sums.push_back(p);
    p += t + 1;
    offs.push_back(p);
    p += t + 1;
  }
  // upsweep
  const uint32_t* cur_in = in;
  uint32_t* cur_out = out;
  size_t cur_n = n;
  for (size_t l = 0; l < sizes.size(); l++) {
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)sizes[l]), dim3(SCAN_THREADS), 0, s, cur_in, cur_out,
                       cur_n, sums[l]);
    if (sizes[l] <= 1) break;
    cur_in = sums[l];
    cur_out = offs[l];
    cur_n = sizes[l];
  }
  // downsweep: level l output gets level l+1's scanned sums added
  for (int l = (int)sizes.size() - 2; l >= 0; l--) {
    uint32_t* target = (l == 0) ? out : offs[l - 1];
    size_t tn = (l == 0) ? n : sizes[l - 1];
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)sizes[l]), dim3(SCAN_THREADS), 0, s, target, tn, offs[l]);
  }
  RTG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- radix --

constexpr int RS_THREADS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint32_t* __restrict__ keys, size_t n, int shift,
                                                        uint32_t* __restrict__ hist, uint32_t num_tiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ROUNDS; r++) {
    size_t pos = base + (size_t)r * RS_THREADS + threadIdx.x;
    if (pos < n) atomicAdd(&h[(keys[pos] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * num_tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals, size_t n,
                                                           int shift, const uint32_t* __restrict__ offs,
                                                           uint32_t num_tiles, uint32_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[RS_WAVES][256];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  run[threadIdx.x] = offs[(size_t)threadIdx.x * num_tiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < RS_WAVES; w++) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const size_t base = (size_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; r++) {
    const size_t pos = base + (size_t)r * RS_THREADS + threadIdx.x;
    const bool valid = pos < n;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = keys[pos];
      val = vals[pos];
      d = (key >> shift) & 255u;
    }
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t wrank = __popcll(peers & lt);
    const uint32_t tag = (uint32_t)(r + 1) << 8;
    if (valid && wrank == 0) wcnt[wave][d] = tag | (uint32_t)__popcll(peers);
    __syncthreads();
    uint32_t prefix = 0;
    if (valid) {
      for (int w = 0; w < wave; w++) {
        uint32_t v = wcnt[w][d];
        prefix += ((v & ~255u) == tag) ? (v & 255u) : 0u;
      }
      const uint32_t dst = run[d] + prefix + wrank;
      if (dst < n) {  // holds for a correct scan; guards memory if it does not
        keys_out[dst] = key;
        vals_out[dst] = val;
      }
    }
    __syncthreads();
    {
      uint32_t add = 0;
#pragma unroll
      for (int w = 0; w < RS_WAVES; w++) {
        uint32_t v = wcnt[w][threadIdx.x];
        add += ((v & ~255u) == tag) ? (v & 255u) : 0u;
      }
      run[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

void RadixSorter::sort_pairs(uint32_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit,
                             hipStream_t s) {
  if (n <= 1) return;
  const uint32_t num_tiles = div_up(n, RS_TILE);
  const size_t hist_n = (size_t)256 * num_tiles;
  hist.reserve(hist_n * sizeof(uint32_t));
  offs.reserve(hist_n * sizeof(uint32_t));
  kt.reserve(n * sizeof(uint32_t));
  vt.reserve(n * sizeof(uint32_t));
  uint32_t *ki = keys, *vi = vals, *ko = kt.as<uint32_t>(), *vo = vt.as<uint32_t>();
  int passes = 0;
  for (int shift = begin_bit; shift < end_bit; shift += 8) {
    hipLaunchKernelGGL(k_rs_hist, dim3(num_tiles), dim3(RS_THREADS), 0, s, ki, n, shift, hist.as<uint32_t>(),
                       num_tiles);
    scanner.exclusive(hist.as<uint32_t>(), offs.as<uint32_t>(), hist_n, s);
    hipLaunchKernelGGL(k_rs_scatter, dim3(num_tiles), dim3(RS_THREADS), 0, s, ki, vi, n, shift,
                       offs.as<uint32_t>(), num_tiles, ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
    passes++;
  }
  if (passes & 1) {
    RTG_HIP(hipMemcpyAsync(keys, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    RTG_HIP(hipMemcpyAsync(vals, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  RTG_HIP(hipGetLastError());
}

}  // namespace rtg
