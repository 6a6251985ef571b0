// Seeded synthetic column generators (device side) and the library runtime
// helpers (launch shape, scratch, last-error text).
//
// Generator: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) used as a counter
// hash of (seed, stream, row), reduced to [lo, hi] with Lemire's
// multiply-shift (TOMACS 2019).  Every value is a pure function of its row, so
// a table of any size is produced in one pass at HBM write speed.
#include "capi_util.hpp"

#include <mutex>
#include <vector>

namespace sdb {
namespace {

constexpr int BLOCK = 256;

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream)
{
    return splitmix64(seed * 0x100000001B3ULL + stream);
}

__device__ __forceinline__ int32_t uniform(uint64_t key, uint64_t row, int32_t lo, uint64_t range)
{
    const uint64_t h = splitmix64(key + row);
    return (int32_t)((int64_t)lo + (int64_t)(((h >> 32) * range) >> 32));
}

__global__ void __launch_bounds__(BLOCK) k_gen_uniform(int32_t *out, int64_t n, uint64_t key, int64_t first,
                                                       int32_t lo, uint64_t range)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = uniform(key, (uint64_t)(first + i), lo, range);
}

__global__ void __launch_bounds__(BLOCK) k_gen_child(int32_t *out, const int32_t *base, int64_t n, uint64_t key,
                                                     int64_t first, int32_t mult, int32_t lo, uint64_t range)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = base[i] * mult + uniform(key, (uint64_t)(first + i), lo, range);
}

__global__ void __launch_bounds__(BLOCK) k_gen_lookup(int32_t *out, int64_t n, uint64_t key, int64_t first,
                                                      const int32_t *table, uint64_t len)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = table[uniform(key, (uint64_t)(first + i), 0, len)];
}

__global__ void __launch_bounds__(BLOCK) k_gen_rowid(int32_t *out, int64_t n, int64_t first)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = (int32_t)(first + i + 1);
}

thread_local std::string g_last_error;

struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_scratch_mu;
std::vector<Scratch> g_scratch;   // per device
std::vector<int> g_cu_count;      // per device, 0 = unknown

} // namespace

void set_last_error(const std::string &msg) { g_last_error = msg; }

int device_cu_count()
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_cu_count.size() <= dev) g_cu_count.resize(dev + 1, 0);
    if (g_cu_count[dev] == 0) {
        int cus = 0;
        SDB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        g_cu_count[dev] = cus > 0 ? cus : 256;
    }
    return g_cu_count[dev];
}

unsigned stream_grid(int64_t units, int64_t per_block, int blocks_per_cu)
{
    if (units <= 0) return 1;
    const int64_t need = (units + per_block - 1) / per_block;
    const int64_t cap = (int64_t)device_cu_count() * blocks_per_cu;
    return (unsigned)std::max<int64_t>(1, std::min(need, cap));
}

void *scratch_buffer(size_t bytes, hipStream_t stream)
{
    int dev = 0;
    SDB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch.size() <= dev) g_scratch.resize(dev + 1);
    Scratch &s = g_scratch[dev];
    if (s.bytes < bytes) {
        if (s.ptr) {
            SDB_HIP(hipStreamSynchronize(stream));
            SDB_HIP(hipFree(s.ptr));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        size_t want = std::max<size_t>(bytes, 1 << 20);
        SDB_HIP(hipMalloc(&s.ptr, want));
        s.bytes = want;
    }
    return s.ptr;
}

} // namespace sdb

using namespace sdb;

extern "C" {

const char *sdb_last_error(void) { return g_last_error.c_str(); }

const char *sdb_version(void) { return "sycldb-mi355x 0.1"; }

int sdb_gen_uniform(int32_t *out, int64_t n, uint64_t seed, uint64_t stream_id, int64_t first_row, int32_t lo,
                    int32_t hi, sdb_stream stream)
{
    return guarded([&] {
        require(n >= 0 && (n == 0 || out), "bad output");
        require(hi >= lo, "hi < lo");
        if (n == 0) return;
        const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
        hipLaunchKernelGGL(k_gen_uniform, dim3(stream_grid(n, BLOCK)), dim3(BLOCK), 0, as_stream(stream), out, n,
                           stream_key(seed, stream_id), first_row, lo, range);
        hip_check(hipGetLastError(), "gen_uniform");
    });
}

int sdb_gen_child(int32_t *out, const int32_t *base, int64_t n, uint64_t seed, uint64_t stream_id,
                  int64_t first_row, int32_t mult, int32_t lo, int32_t hi, sdb_stream stream)
{
    return guarded([&] {
        require(n >= 0 && (n == 0 || (out && base)), "bad pointers");
        require(hi >= lo, "hi < lo");
        if (n == 0) return;
        const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
        hipLaunchKernelGGL(k_gen_child, dim3(stream_grid(n, BLOCK)), dim3(BLOCK), 0, as_stream(stream), out, base,
                           n, stream_key(seed, stream_id), first_row, mult, lo, range);
        hip_check(hipGetLastError(), "gen_child");
    });
}

int sdb_gen_lookup(int32_t *out, int64_t n, uint64_t seed, uint64_t stream_id, int64_t first_row,
                   const int32_t *table, int32_t table_len, sdb_stream stream)
{
    return guarded([&] {
        require(n >= 0 && (n == 0 || (out && table)), "bad pointers");
        require(table_len > 0, "empty lookup table");
        if (n == 0) return;
        hipLaunchKernelGGL(k_gen_lookup, dim3(stream_grid(n, BLOCK)), dim3(BLOCK), 0, as_stream(stream), out, n,
                           stream_key(seed, stream_id), first_row, table, (uint64_t)table_len);
        hip_check(hipGetLastError(), "gen_lookup");
    });
}

int sdb_device_alloc(void **ptr, uint64_t bytes)
{
    return guarded([&] {
        require(ptr != nullptr, "null output pointer");
        *ptr = nullptr;
        if (bytes == 0) return;
        const hipError_t e = hipMalloc(ptr, bytes);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            fail(SDB_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
        }
        hip_check(e, "hipMalloc");
    });
}

int sdb_device_free(void *ptr)
{
    return guarded([&] {
        if (ptr) SDB_HIP(hipFree(ptr));
    });
}

int sdb_copy_to_device(void *dst, const void *host_src, uint64_t bytes, sdb_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        require(dst && host_src, "null pointer");
        SDB_HIP(hipMemcpyAsync(dst, host_src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
        SDB_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int sdb_copy_to_host(void *host_dst, const void *src, uint64_t bytes, sdb_stream stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        require(host_dst && src, "null pointer");
        SDB_HIP(hipMemcpyAsync(host_dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
        SDB_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}

int sdb_stream_synchronize(sdb_stream stream)
{
    return guarded([&] { SDB_HIP(hipStreamSynchronize(as_stream(stream))); });
}

int sdb_gen_rowid(int32_t *out, int64_t n, int64_t first_row, sdb_stream stream)
{
    return guarded([&] {
        require(n >= 0 && (n == 0 || out), "bad output");
        if (n == 0) return;
        hipLaunchKernelGGL(k_gen_rowid, dim3(stream_grid(n, BLOCK)), dim3(BLOCK), 0, as_stream(stream), out, n,
                           first_row);
        hip_check(hipGetLastError(), "gen_rowid");
    });
}

} // extern "C"
