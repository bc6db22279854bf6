// Device sort and scan of libssq: a stable LSD radix sort of (key, int32 value) pairs with 8-bit
// digits, and an exclusive int64 scan.  Wave64 throughout; no library kernels.
//
// Radix sort, per pass over one digit:
//  * k_rs_hist: each workgroup counts the digits of its tile of 4096 keys in LDS and writes the
//    256 counts digit-major (hist[digit * tiles + tile]);
//  * the digit-major counts are scanned, which gives every (digit, tile) its first output slot;
//  * k_rs_scatter: the tile is placed in 16 rounds of 256 keys in index order; within a round a
//    key's rank among equal digits comes from 8 ballots (the lanes of its wavefront with the
//    same digit) and the counts of the wavefronts before it, so equal digits keep their input
//    order (stable).
// The passes ping-pong between the output and a scratch copy so that the last pass writes the
// output; the input is not modified.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace ssq_sort {
namespace {  // one copy per translation unit

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;  // keys per workgroup and pass
constexpr int kRadix = 256;
constexpr int kWaves = kThreads / 64;

template <typename Key>
__global__ void __launch_bounds__(kThreads) k_rs_hist(const Key* __restrict__ keys, int64_t n, int shift,
                                                      int64_t tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[kRadix];
  cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * kTile;
  for (int r = 0; r < kItems; r++) {
    const int64_t i = t0 + (int64_t)r * kThreads + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

template <typename Key>
__global__ void __launch_bounds__(kThreads) k_rs_scatter(const Key* __restrict__ kin, const int32_t* __restrict__ vin,
                                                         Key* __restrict__ kout, int32_t* __restrict__ vout, int64_t n,
                                                         int shift, int64_t tiles, const uint32_t* __restrict__ offs) {
  __shared__ uint32_t base[kRadix];        // next output slot of each digit for this tile
  __shared__ uint32_t wc[kWaves][kRadix];  // this round's digit counts per wavefront
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  base[tid] = offs[(int64_t)tid * tiles + blockIdx.x];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const int64_t t0 = (int64_t)blockIdx.x * kTile;
  for (int r = 0; r < kItems; r++) {
    for (int w = 0; w < kWaves; w++) wc[w][tid] = 0u;
    __syncthreads();
    const int64_t i = t0 + (int64_t)r * kThreads + tid;
    const bool valid = i < n;
    Key k = 0;
    int32_t v = 0;
    uint32_t d = 0;
    if (valid) {
      k = kin[i];
      v = vin[i];
      d = (uint32_t)(k >> shift) & 255u;
    }
    uint64_t eq = __ballot(valid);  // valid lanes with this lane's digit
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(valid && bit);
      eq &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(eq & below);
    if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(eq);
    __syncthreads();
    if (valid) {
      uint32_t before = 0;
      for (int w = 0; w < wave; w++) before += wc[w][d];
      const uint32_t pos = base[d] + before + rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kWaves; w++) add += wc[w][tid];
    base[tid] += add;
    __syncthreads();
  }
}

// ---- prefix sums, three passes: tile partials, tile sums, tile offsets -----------------------
// Each workgroup scans kTile consecutive values (16 per thread, then the 256 thread sums); the
// sums are carried in int64.
template <typename In, typename Out, bool Inclusive>
__global__ void __launch_bounds__(kThreads) k_sc_partial(const In* __restrict__ in, int64_t n,
                                                         Out* __restrict__ out, int64_t* __restrict__ tsum) {
  __shared__ int64_t s[kThreads];
  const int tid = threadIdx.x;
  const int64_t b0 = (int64_t)blockIdx.x * kTile + (int64_t)tid * kItems;
  int64_t v[kItems];
  int64_t acc = 0;
  for (int k = 0; k < kItems; k++) {
    const int64_t i = b0 + k;
    v[k] = i < n ? (int64_t)in[i] : 0;
    acc += v[k];
  }
  s[tid] = acc;
  __syncthreads();
  for (int o = 1; o < kThreads; o <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
    const int64_t x = tid >= o ? s[tid - o] : 0;
    __syncthreads();
    s[tid] += x;
    __syncthreads();
  }
  int64_t run = s[tid] - acc;  // exclusive start of this thread's values
  for (int k = 0; k < kItems; k++) {
    const int64_t i = b0 + k;
    if (Inclusive) run += v[k];
    if (i < n) out[i] = (Out)run;
    if (!Inclusive) run += v[k];
  }
  if (tid == kThreads - 1) tsum[blockIdx.x] = s[tid];
}

// one workgroup: exclusive scan of the nt tile sums in place; tsum[nt] = the total
__global__ void __launch_bounds__(kThreads) k_sc_tiles(int64_t nt, int64_t* __restrict__ tsum) {
  __shared__ int64_t s[kThreads];
  __shared__ int64_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t c0 = 0; c0 < nt; c0 += kThreads) {
    const int64_t i = c0 + tid;
    const int64_t x0 = i < nt ? tsum[i] : 0;
    s[tid] = x0;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
      const int64_t x = tid >= o ? s[tid - o] : 0;
      __syncthreads();
      s[tid] += x;
      __syncthreads();
    }
    const int64_t base = carry;
    if (i < nt) tsum[i] = base + s[tid] - x0;
    __syncthreads();
    if (tid == kThreads - 1) carry = base + s[tid];
    __syncthreads();
  }
  if (tid == 0) tsum[nt] = carry;
}

template <typename Out>
__global__ void __launch_bounds__(kThreads) k_sc_add(Out* __restrict__ out, int64_t n,
                                                     const int64_t* __restrict__ tsum) {
  const int64_t b0 = (int64_t)blockIdx.x * kTile;
  const Out add = (Out)tsum[blockIdx.x];
  for (int k = 0; k < kItems; k++) {
    const int64_t i = b0 + (int64_t)k * kThreads + threadIdx.x;
    if (i < n) out[i] += add;
  }
}

template <typename Out>
__global__ void k_sc_total(const int64_t* __restrict__ tsum, int64_t nt, Out* __restrict__ out_n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *out_n = (Out)tsum[nt];
}

// uint32 digit counts scanned through the int64 scan (counts, then slots back to uint32)
__global__ void k_rs_narrow(const int64_t* __restrict__ in, int64_t m, uint32_t* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)in[i];
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline int64_t tiles_of(int64_t n) { return n > 0 ? (n + kTile - 1) / kTile : 1; }

// Scratch bytes of scan_exclusive for n values.
inline size_t scan_bytes(int64_t n) { return align256(sizeof(int64_t) * (size_t)(tiles_of(n) + 1)); }

// Inclusive (out[i] = in[0] + ... + in[i]) or exclusive prefix sums of in[0, n); the exclusive
// form also writes out[n] = the total.  Scratch of scan_bytes(n) bytes.
template <typename In, typename Out, bool Inclusive>
hipError_t scan(const In* in, int64_t n, Out* out, void* scratch, hipStream_t s) {
  const int64_t nt = tiles_of(n);
  int64_t* tsum = (int64_t*)scratch;
  if (n > 0) {
    hipLaunchKernelGGL((k_sc_partial<In, Out, Inclusive>), dim3((unsigned)nt), dim3(kThreads), 0, s, in, n, out, tsum);
    hipLaunchKernelGGL(k_sc_tiles, dim3(1), dim3(kThreads), 0, s, nt, tsum);
    hipLaunchKernelGGL(k_sc_add<Out>, dim3((unsigned)nt), dim3(kThreads), 0, s, out, n, (const int64_t*)tsum);
  } else {
    hipLaunchKernelGGL(k_sc_tiles, dim3(1), dim3(kThreads), 0, s, (int64_t)0, tsum);  // tsum[0] = 0
  }
  if (!Inclusive) hipLaunchKernelGGL(k_sc_total<Out>, dim3(1), dim3(1), 0, s, (const int64_t*)tsum, n > 0 ? nt : 0, out + n);
  return hipGetLastError();
}

// Stable sort of (kin, vin) by key into (kout, vout), for keys below 2^end_bit: the passes
// cover whole 8-bit digits, ceil(end_bit / 8) of them, so any bits of a key up to the next
// multiple of 8 take part.  With tmp == nullptr, only sets bytes (the scratch size); count < 2^32.
template <typename Key>
hipError_t sort_pairs(void* tmp, size_t& bytes, const Key* kin, Key* kout, const int32_t* vin, int32_t* vout,
                      int64_t count, int end_bit, hipStream_t s) {
  const int64_t tiles = tiles_of(count), m = (int64_t)kRadix * tiles;
  const size_t b_keys = align256(sizeof(Key) * (size_t)count), b_vals = align256(sizeof(int32_t) * (size_t)count),
               b_h32 = align256(sizeof(uint32_t) * (size_t)m), b_h64 = align256(sizeof(int64_t) * (size_t)(m + 1));
  const size_t need = b_keys + b_vals + b_h32 + b_h64 + scan_bytes(m);
  if (!tmp) {
    bytes = need;
    return hipSuccess;
  }
  if (bytes < need) return hipErrorInvalidValue;
  if (count <= 0) return hipSuccess;
  char* p = (char*)tmp;
  Key* kt = (Key*)p;
  int32_t* vt = (int32_t*)(p + b_keys);
  uint32_t* h32 = (uint32_t*)(p + b_keys + b_vals);
  int64_t* h64 = (int64_t*)(p + b_keys + b_vals + b_h32);
  void* sc = p + b_keys + b_vals + b_h32 + b_h64;
  const int passes = end_bit > 0 ? (end_bit + 7) / 8 : 1;
  const Key* ks = kin;
  const int32_t* vs = vin;
  for (int q = 0; q < passes; q++) {
    const bool to_out = ((passes - 1 - q) & 1) == 0;  // the last pass writes the output
    Key* kd = to_out ? kout : kt;
    int32_t* vd = to_out ? vout : vt;
    const int shift = 8 * q;
    hipLaunchKernelGGL(k_rs_hist<Key>, dim3((unsigned)tiles), dim3(kThreads), 0, s, ks, count, shift, tiles, h32);
    hipError_t e = scan<uint32_t, int64_t, false>(h32, m, h64, sc, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_narrow, dim3((unsigned)std::min<int64_t>((m + 255) / 256, 4096)), dim3(256), 0, s,
                       (const int64_t*)h64, m, h32);
    hipLaunchKernelGGL(k_rs_scatter<Key>, dim3((unsigned)tiles), dim3(kThreads), 0, s, ks, vs, kd, vd, count, shift,
                       tiles, (const uint32_t*)h32);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    ks = kd;
    vs = vd;
  }
  return hipSuccess;
}

}  // namespace
}  // namespace ssq_sort
