// Residency probe (test infrastructure): how many blocks of a given shape run
// on one compute unit at the same time. Every block records its CU (HW_ID,
// XCC_ID) and its start and end times (s_memrealtime) and spins 200 us; no
// block waits on another, so blocks that do not fit simply run later, and the
// largest number of overlapping blocks on a CU is its residency. The kernel's
// VGPRs and SGPRs are fixed by register clobbers (the compiler allocates up to
// the highest register named), its LDS by the launch.
//
// Output, one line per shape:
//   <name> threads <t> vgpr <v> sgpr_clobber <s> lds <bytes> query <blocks per CU
//   from hipOccupancyMaxActiveBlocksPerMultiprocessor> resident_min <m> resident_max <M> cus <n>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <map>
#include <vector>

#define HW_REG_HW_ID 4
#define HW_REG_XCC_ID 20
#define GETREG(id) __builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | (id))

template <int VG, int SG>
__global__ void __launch_bounds__(1024) k_resid(unsigned long long *rec, unsigned long long spin,
                                                 int lds_ints) {
  extern __shared__ int dyn[];
  __shared__ int st[50];
  if constexpr (VG == 62) asm volatile("s_nop 0" ::: "v61");
  if constexpr (VG == 114) asm volatile("s_nop 0" ::: "v113");
  if constexpr (SG == 84) asm volatile("s_nop 0" ::: "s83");
  if constexpr (SG == 100) asm volatile("s_nop 0" ::: "s99");
  if (threadIdx.x < 50) st[threadIdx.x] = (int)threadIdx.x;
  if ((int)threadIdx.x < lds_ints) dyn[threadIdx.x] = (int)threadIdx.x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned hwid = GETREG(HW_REG_HW_ID);
    const unsigned xcc = GETREG(HW_REG_XCC_ID) & 15u;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    unsigned long long t = t0;
    for (int it = 0; t - t0 < spin && it < 4000000; ++it) {
      __builtin_amdgcn_s_sleep(1);
      t = __builtin_amdgcn_s_memrealtime();
    }
    rec[blockIdx.x * 4 + 0] = hwid;
    rec[blockIdx.x * 4 + 1] = xcc;
    rec[blockIdx.x * 4 + 2] = t0;
    rec[blockIdx.x * 4 + 3] = t + (unsigned long long)(st[1] - 1);
  }
  __syncthreads();
}

#define CHECK(x)                                                              \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess) {                                                   \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                 \
      exit(1);                                                                \
    }                                                                         \
  } while (0)

template <int VG, int SG>
static void run(const char *name, int threads, int blocks, int dyn_lds) {
  unsigned long long *rec = nullptr;
  CHECK(hipMalloc(&rec, sizeof(unsigned long long) * 4 * blocks));
  CHECK(hipMemset(rec, 0, sizeof(unsigned long long) * 4 * blocks));
  int query = 0;
  CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&query, k_resid<VG, SG>, threads, dyn_lds));
  hipFuncAttributes fa;
  CHECK(hipFuncGetAttributes(&fa, (const void *)k_resid<VG, SG>));
  // s_memrealtime runs at 100 MHz: 20,000 ticks = 200 us
  hipLaunchKernelGGL((k_resid<VG, SG>), dim3(blocks), dim3(threads), dyn_lds, 0, rec, 20000ull,
                     dyn_lds / 4);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  std::vector<unsigned long long> h(4 * (size_t)blocks);
  CHECK(hipMemcpy(h.data(), rec, sizeof(unsigned long long) * 4 * blocks, hipMemcpyDeviceToHost));
  CHECK(hipFree(rec));
  // CU key: XCC, shader engine [15:13], shader array [12], CU [11:8] of HW_ID
  std::map<unsigned, std::vector<std::pair<unsigned long long, int>>> ev;
  for (int b = 0; b < blocks; ++b) {
    const unsigned hw = (unsigned)h[4 * b], xcc = (unsigned)h[4 * b + 1];
    const unsigned key = (xcc << 16) | (((hw >> 13) & 7u) << 8) | (((hw >> 12) & 1u) << 4) |
                         ((hw >> 8) & 15u);
    ev[key].push_back({h[4 * b + 2], +1});
    ev[key].push_back({h[4 * b + 3], -1});
  }
  int mx = 0, mn = 1 << 30;
  for (auto &kv : ev) {
    auto v = kv.second;
    std::sort(v.begin(), v.end(), [](const std::pair<unsigned long long, int> &a,
                                     const std::pair<unsigned long long, int> &b) {
      return a.first < b.first || (a.first == b.first && a.second < b.second);
    });
    int c = 0, m = 0;
    for (auto &p : v) {
      c += p.second;
      m = std::max(m, c);
    }
    mx = std::max(mx, m);
    mn = std::min(mn, m);
  }
  printf("%s threads %d vgpr %d sgpr_clobber %d lds %zu query %d resident_min %d resident_max %d cus %zu\n",
         name, threads, fa.numRegs, SG, fa.sharedSizeBytes + (size_t)dyn_lds, query, mn, mx,
         ev.size());
  fflush(stdout);
}

int main() {
  int ncu = 0;
  CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  const int B = ncu * 12;  // more blocks than can be resident
  run<0, 0>("light_256", 256, B, 8000);
  run<62, 0>("vgpr62_256", 256, B, 8000);
  run<0, 84>("sgpr84_256", 256, B, 8000);
  run<0, 100>("sgpr100_256", 256, B, 8000);
  run<62, 84>("vgpr62_sgpr84_256", 256, B, 8000);
  run<62, 84>("vgpr62_sgpr84_256_nodyn", 256, B, 0);
  run<114, 0>("vgpr114_256", 256, B, 8000);
  run<114, 0>("vgpr114_512", 512, B, 8000);
  run<0, 84>("sgpr84_64", 64, ncu * 40, 0);
  return 0;
}
