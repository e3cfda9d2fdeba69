// Seeded uniform phase masks for synthetic request streams.
//
// Counter-based: element i of member b is 2 pi u - pi with u in [0, 1) taken
// from the splitmix64 finaliser of seed_b + (i + 1) * golden-ratio increment,
// so any element of any member is generated independently of the others and
// a captured graph regenerates a batch from device-held seeds.  (No
// counterpart in the reference, whose phases come from the caller.)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xl {

struct PhaseArgs {
  float* out;
  const uint64_t* seeds;   // one per member
  const int32_t* enable;   // member b is written when enable[b] != 0 (NULL: all)
  int64_t count;           // elements per member
  int64_t bs;              // batch stride (elements)
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// u in [0, 1) from the top 24 bits (exact in float32)
__device__ __forceinline__ float unit_from_bits(uint64_t h) {
  return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f);
}

__global__ void uniform_phase_kernel(PhaseArgs a) {
  const int b = blockIdx.y;
  if (a.enable && !a.enable[b]) return;
  const uint64_t key = a.seeds[b] * 0xD1B54A32D192ED03ull;
  float* o = a.out + b * a.bs;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < a.count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float u = unit_from_bits(splitmix64(key + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull));
    o[i] = fmaf(6.28318530717958648f, u, -3.14159265358979324f);
  }
}

}  // namespace xl
