// Test harness of libssq's device sort and scans (specseqplusplus_amd/csrc/ssq_sort.h): host
// buffers in, host buffers out, so that GPU tests can check the stable LSD radix sort against a
// stable host sort and the scans against host prefix sums.  Test infrastructure only: the
// product calls ssq_sort.h from ssq_ripsco.hip, never this library.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../specseqplusplus_amd/csrc/ssq_sort.h"

namespace {

constexpr int kOk = 0, kInvalid = 1, kDevice = 2;

// device buffers freed on every return path
struct Dev {
  std::vector<void*> ptrs;
  ~Dev() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(int64_t count) {
    void* p = nullptr;
    const size_t bytes = sizeof(T) * (size_t)(count > 0 ? count : 1);
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
};

bool up(void* d, const void* h, size_t bytes) {
  return bytes == 0 || hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) == hipSuccess;
}
bool down(void* h, const void* d, size_t bytes) {
  return bytes == 0 || hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) == hipSuccess;
}

template <typename In, typename Out, bool Inclusive>
int run_scan(const In* in, int64_t n, int64_t* out) {
  Dev D;
  In* din = D.alloc<In>(n);
  Out* dout = D.alloc<Out>(n + 1);
  uint8_t* scratch = D.alloc<uint8_t>((int64_t)ssq_sort::scan_bytes(n));
  if (!din || !dout || !scratch) return kDevice;
  if (!up(din, in, sizeof(In) * (size_t)n)) return kDevice;
  if (ssq_sort::scan<In, Out, Inclusive>(din, n, dout, scratch, nullptr) != hipSuccess) return kDevice;
  const int64_t m = Inclusive ? n : n + 1;
  std::vector<Out> h((size_t)(m > 0 ? m : 1));
  if (hipDeviceSynchronize() != hipSuccess || !down(h.data(), dout, sizeof(Out) * (size_t)m)) return kDevice;
  for (int64_t i = 0; i < m; i++) out[i] = (int64_t)h[(size_t)i];
  return kOk;
}

}  // namespace

// Stable sort of (keys, vals) by key bits [0, end_bit) (1 <= end_bit <= 64), as ssq_ripsco.hip
// sorts its edge and tie keys.
extern "C" int ssq_harness_sort_u64(const uint64_t* keys, const int32_t* vals, int64_t n, int end_bit,
                                    uint64_t* kout, int32_t* vout) {
  if (n < 0 || n >= (int64_t(1) << 32) || end_bit < 1 || end_bit > 64 ||
      (n > 0 && (!keys || !vals || !kout || !vout)))
    return kInvalid;
  Dev D;
  uint64_t* dk = D.alloc<uint64_t>(n);
  uint64_t* dk2 = D.alloc<uint64_t>(n);
  int32_t* dv = D.alloc<int32_t>(n);
  int32_t* dv2 = D.alloc<int32_t>(n);
  if (!dk || !dk2 || !dv || !dv2) return kDevice;
  if (!up(dk, keys, sizeof(uint64_t) * (size_t)n) || !up(dv, vals, sizeof(int32_t) * (size_t)n)) return kDevice;
  size_t bytes = 0;
  if (ssq_sort::sort_pairs<uint64_t>(nullptr, bytes, dk, dk2, dv, dv2, n, end_bit, nullptr) != hipSuccess)
    return kDevice;
  uint8_t* tmp = D.alloc<uint8_t>((int64_t)bytes);
  if (!tmp) return kDevice;
  if (ssq_sort::sort_pairs<uint64_t>(tmp, bytes, dk, dk2, dv, dv2, n, end_bit, nullptr) != hipSuccess)
    return kDevice;
  if (hipDeviceSynchronize() != hipSuccess || !down(kout, dk2, sizeof(uint64_t) * (size_t)n) ||
      !down(vout, dv2, sizeof(int32_t) * (size_t)n))
    return kDevice;
  return kOk;
}

// Prefix sums of n values into out (int64 on the host):
//   mode 0: exclusive, int32 in, int64 out, n + 1 values (out[n] = the total);
//   mode 1: exclusive, uint32 in, int64 out, n + 1 values (the radix sort's digit counts);
//   mode 2: inclusive, int32 in, int32 out, n values.
extern "C" int ssq_harness_scan(const void* in, int64_t n, int mode, int64_t* out) {
  if (n < 0 || n >= (int64_t(1) << 40) || !out || (n > 0 && !in)) return kInvalid;
  switch (mode) {
    case 0:
      return run_scan<int32_t, int64_t, false>((const int32_t*)in, n, out);
    case 1:
      return run_scan<uint32_t, int64_t, false>((const uint32_t*)in, n, out);
    case 2:
      return run_scan<int32_t, int32_t, true>((const int32_t*)in, n, out);
    default:
      return kInvalid;
  }
}
