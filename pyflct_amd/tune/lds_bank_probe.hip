// LDS bank-conflict probe for gfx950.
//
// Each launch reads one lane -> address pattern from LDS kIter times per
// wave (one wave per workgroup, 4096 workgroups); rocprofv3 counts
// SQ_LDS_BANK_CONFLICT and SQ_LDS_IDX_ACTIVE per dispatch, and
// lds_bank_table.py turns them into cycles per load.  The patterns are the
// fragment loads of the split-BF16 kernels: DFT table rows (ds_read_b128 of
// row lane & 15 at dword 4 (lane >> 4)) and box-row spectrum reads
// (ds_read2_b32 / ds_read_b64 at a row stride of a dwords), for a range of
// row strides.  Every address is checked on the host to lie inside the
// 64 KB array before the launch.
//
// Build: hipcc -O2 --offload-arch=gfx950 -Xclang -target-feature -Xclang
//        -packed-fp32-ops -o lds_bank_probe lds_bank_probe.hip
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

constexpr int kDw = 16384;  // dwords of LDS (64 KB)
constexpr int kIter = 256;
constexpr int kGrid = 4096;

#define CHECK(x)                                                     \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));        \
      exit(1);                                                       \
    }                                                                \
  } while (0)

// kind 0: ds_read2_b32 (dwords d, d+1); 1: ds_read_b128 (d..d+3, d % 4 == 0);
// 2: ds_read_b64 (d, d+1, d even); 3: ds_read_b32.  Each asm block waits
// for its own load.
__global__ void __launch_bounds__(64) k_probe(const int* __restrict__ addr, int kind, unsigned* out) {
  __shared__ unsigned lds[kDw];
  for (int i = threadIdx.x; i < kDw; i += 64) lds[i] = i * 2654435761u;
  // the array's LDS address, handed to the asm blocks (keeps the array and
  // its stores)
  const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) unsigned*)lds;
  asm volatile("" ::"v"(base) : "memory");
  __syncthreads();
  const unsigned ba = base + (unsigned)addr[threadIdx.x] * 4u;
  unsigned acc = 0;
  for (int it = 0; it < kIter; ++it) {
    if (kind == 0) {
      unsigned long long r;
      asm volatile("ds_read2_b32 %0, %1 offset1:1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(ba) : "memory");
      acc += (unsigned)r + (unsigned)(r >> 32);
    } else if (kind == 1) {
      uint4 r;
      asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(ba) : "memory");
      acc += r.x + r.y + r.z + r.w;
    } else if (kind == 2) {
      unsigned long long r;
      asm volatile("ds_read_b64 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(ba) : "memory");
      acc += (unsigned)r + (unsigned)(r >> 32);
    } else {
      unsigned r;
      asm volatile("ds_read_b32 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(r) : "v"(ba) : "memory");
      acc += r;
    }
  }
  out[blockIdx.x * 64 + threadIdx.x] = acc;
}

struct Pat {
  char name[64];
  int kind;
  int d[64];
};

template <class F>
static Pat make(int kind, const char* name, F f) {
  Pat p;
  snprintf(p.name, sizeof p.name, "%s", name);
  p.kind = kind;
  for (int l = 0; l < 64; ++l) p.d[l] = f(l);
  return p;
}

int main() {
  std::vector<Pat> pats;
  char nm[64];
  pats.push_back(make(3, "b32 consecutive", [](int l) { return l; }));
  pats.push_back(make(3, "b32 stride 64 dwords", [](int l) { return 64 * l; }));
  pats.push_back(make(3, "b32 stride 32 dwords", [](int l) { return 32 * l; }));
  pats.push_back(make(0, "read2_b32 consecutive pairs", [](int l) { return 2 * l; }));
  pats.push_back(make(2, "b64 consecutive", [](int l) { return 2 * l; }));
  pats.push_back(make(1, "b128 consecutive", [](int l) { return 4 * l; }));
  for (int a : {12, 16, 20, 24, 28, 32, 36, 40, 44}) {
    snprintf(nm, sizeof nm, "b128 table rows a=%d", a);
    pats.push_back(make(1, nm, [a](int l) { return a * (l & 15) + 4 * (l >> 4); }));
  }
  for (int a : {130, 132, 134, 136, 138, 140, 144, 146, 160}) {
    snprintf(nm, sizeof nm, "read2_b32 box rows a=%d", a);
    pats.push_back(make(0, nm, [a](int l) { return a * (l & 15) + 4 * (l >> 4); }));
  }
  for (int a : {134, 136, 138, 144}) {
    snprintf(nm, sizeof nm, "b64 box rows a=%d", a);
    pats.push_back(make(2, nm, [a](int l) { return a * (l & 15) + 4 * (l >> 4); }));
  }
  int* daddr = nullptr;
  unsigned* dout = nullptr;
  CHECK(hipMalloc(&daddr, 64 * sizeof(int)));
  CHECK(hipMalloc(&dout, (size_t)kGrid * 64 * sizeof(unsigned)));
  for (const Pat& p : pats) {
    const int span = (p.kind == 1) ? 4 : (p.kind == 3 ? 1 : 2);
    for (int l = 0; l < 64; ++l) {
      const int d = p.d[l];
      if (d < 0 || d + span > kDw || (p.kind == 1 && d % 4) || (p.kind == 2 && d % 2)) {
        fprintf(stderr, "pattern %s: lane %d address %d out of range or misaligned\n", p.name, l, d);
        return 1;
      }
    }
    CHECK(hipMemcpy(daddr, p.d, sizeof p.d, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe, dim3(kGrid), dim3(64), 0, 0, daddr, p.kind, dout);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    printf("pattern: %s\n", p.name);
    fflush(stdout);
  }
  CHECK(hipFree(daddr));
  CHECK(hipFree(dout));
  return 0;
}
