// RadixPartition for callers that cannot allocate device memory (code under HIP graph capture,
// serving loops that forbid hipMalloc): the temporaries of qf_radix_partition (per-tile digit
// histograms, their scan sums, the intermediate pairs of a two-pass digit) come from one
// file-scope device array instead of the context pool.
//
// The array is a fixed part of the code object: every GPU the library runs on holds its own copy
// (36 MiB) until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device, looked up once per device when a context
// is created there.  Every call on one GPU uses the same array and leaves its contents behind, so
// calls that overlap on one GPU (two contexts or streams) race.
#include "qf_internal.h"

#include <mutex>

namespace qf {

// Largest input: the histograms and scan sums of two 8-bit passes over 1,024 tiles of 4,096 rows
// (2 MiB + 1 KiB each) and one set of intermediate (value, row id) columns (2 x 16 MiB).
constexpr int64_t NOALLOC_MAX_ROWS = QF_NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_SCRATCH_BYTES = (36u << 20) + (64u << 10);

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];
static hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];

// Looks up this device's copy of the array (the current device must be `device`); called by
// qf_context_create, and again by the entry point for a device whose lookup has not happened.
void noalloc_scratch_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = ptr;
  else (void)hipGetLastError();
}

static void* noalloc_scratch(int device) {
  QF_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, "device index out of range for the scratch table");
  noalloc_scratch_setup(device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device])
    throw Error(QF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[device]));
  return g_scratch_ptr[device];
}

}  // namespace qf

using namespace qf;

extern "C" int qf_radix_partition_noalloc(qf_context* c, const int32_t* d_values, int64_t n, int radix_bits,
                                          int32_t* d_out_pairs, int64_t* d_offsets) {
  return guarded(c, [&](Context* ctx) {
    QF_REQUIRE(radix_bits >= 1 && radix_bits <= 16, "radix_bits must be in [1, 16]");
    QF_REQUIRE(n >= 0, "row count out of range");
    QF_REQUIRE(n <= NOALLOC_MAX_ROWS, "qf_radix_partition_noalloc takes at most QF_NOALLOC_MAX_ROWS (4,194,304) rows");
    QF_REQUIRE(n == 0 || (d_values && d_out_pairs), "null column");
    QF_REQUIRE(d_offsets, "null offsets");
    Scratch s;
    s.base = static_cast<char*>(noalloc_scratch(ctx->device));
    s.capacity = NOALLOC_SCRATCH_BYTES;
    QF_REQUIRE(partition_scratch_bytes(n, 2, radix_bits) <= s.capacity, "scratch region too small for this input");
    const int32_t* in[1] = {d_values};
    int32_t* out[2] = {d_out_pairs, d_out_pairs + 1};
    stable_partition(ctx, in, 1, 0, true, n, 0, radix_bits, out, 2, d_offsets, &s);
  });
}
