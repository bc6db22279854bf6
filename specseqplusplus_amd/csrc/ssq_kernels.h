// Device-side building blocks shared by the libssq HIP translation units (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssq {

constexpr int kWave = 64;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // empty pivot-table entry (unsigned max for atomicMin)

enum ColState : uint8_t { ST_PENDING = 0, ST_PIVOT = 1, ST_ZERO = 2 };

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Lanes of one wavefront exchange data through LDS; this orders those accesses.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    int t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ int wave_bcast(int v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ long long wave_bcast64(long long v, int src) { return __shfl(v, src, kWave); }

// Append the lanes with pred set to list (ordered within the wave), one atomic per wave.
__device__ __forceinline__ void wave_push(bool pred, int32_t value, int32_t* list, uint32_t* count) {
  const uint64_t mask = __ballot(pred);
  if (!mask) return;
  const int lane = lane_id();
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, leader, kWave);
  if (pred) {
    const uint64_t lower = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    list[base + __popcll(lower)] = value;
  }
}

// Merge-path split of sorted sets A[0:a), B[0:b) at merged position k (ties: A first),
// moved past an equal pair that it would separate so that x in A and x in B always fall
// in the same segment.  Returns i (elements of A before the split); *j gets B's count.
__device__ __forceinline__ int split_adj(const int32_t* A, int a, const int32_t* B, int b, int k,
                                         int* j_out) {
  int lo = k - b > 0 ? k - b : 0;
  int hi = k < a ? k : a;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A[mid] <= B[k - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  int j = k - lo;
  if (lo > 0 && j < b && A[lo - 1] == B[j]) j++;
  *j_out = j;
  return lo;
}

// Symmetric difference of the segment A[i:ie) x B[j:je); counts (emit=false) or writes.
template <bool kEmit>
__device__ __forceinline__ int seg_xor(const int32_t* A, int i, int ie, const int32_t* B, int j, int je,
                                       int32_t* out) {
  int c = 0;
  while (i < ie && j < je) {
    const int32_t x = A[i], y = B[j];
    if (x < y) {
      if (kEmit) out[c] = x;
      c++;
      i++;
    } else if (y < x) {
      if (kEmit) out[c] = y;
      c++;
      j++;
    } else {
      i++;
      j++;
    }
  }
  for (; i < ie; i++, c++)
    if (kEmit) out[c] = A[i];
  for (; j < je; j++, c++)
    if (kEmit) out[c] = B[j];
  return c;
}

// One wavefront: O = A xor B for sorted sets staged in LDS.  Returns |O|.
__device__ __forceinline__ int wave_xor_lds(const int32_t* A, int a, const int32_t* B, int b, int32_t* O) {
  const int lane = lane_id();
  const int total = a + b;
  const int seg = (total + kWave - 1) / kWave;
  int k0 = lane * seg;
  if (k0 > total) k0 = total;
  // a lane's segment ends where the next lane's begins (k1 of lane i is k0 of lane i + 1, and
  // the split is a function of the position): one search per lane, the end from the neighbour
  int j0;
  const int i0 = split_adj(A, a, B, b, k0, &j0);
  int i1 = __shfl_down(i0, 1, kWave), j1 = __shfl_down(j0, 1, kWave);
  if (lane == kWave - 1) {  // k1 == total here: the split after everything
    i1 = a;
    j1 = b;
  }
  const int cnt = seg_xor<false>(A, i0, i1, B, j0, j1, nullptr);
  // each lane's output offset and the total, by bit planes of the counts (cnt <= seg, and seg
  // is the same in every lane): one ballot per bit, the lanes below counted by mbcnt
  int excl = 0, r = 0;
  const int nbits = seg > 0 ? 32 - __clz(seg) : 0;
  for (int bit = 0; bit < nbits; bit++) {
    const uint64_t m = __ballot((cnt >> bit) & 1);
    const int below = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    excl += below << bit;
    r += __popcll(m) << bit;
  }
  seg_xor<true>(A, i0, i1, B, j0, j1, O + excl);
  wave_sync();
  return r;
}

// Rips edge length: sqrt of the float64 sum of squared coordinate differences, summed in
// coordinate order and rounded once to float32 (the host builder's definition, ssq_gen.cpp).
__device__ __forceinline__ float rips_dist(const float* P, int64_t i, int64_t j, int pdim) {
  double acc = 0.0;
  for (int k = 0; k < pdim; k++) {
    const double d = (double)P[i * pdim + k] - (double)P[j * pdim + k];
    acc += d * d;
  }
  return (float)sqrt(acc);
}

__device__ __forceinline__ void wave_copy(int32_t* dst, const int32_t* src, int n) {
  for (int t = lane_id(); t < n; t += kWave) dst[t] = src[t];
}

}  // namespace ssq
