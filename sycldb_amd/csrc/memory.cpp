// Device bump allocator, the MI355X counterpart of
// operations/memory_manager.{hpp,cpp}.
#include <new>

#include "engine.hpp"

namespace sdb {

namespace {
constexpr uint64_t ALIGN = 256;
uint64_t round_up(uint64_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }
} // namespace

// zero budget = budget / 64 as in memory_manager.cpp:55-69
MemoryManager::MemoryManager(uint64_t budget, hipStream_t stream, AllocRegistry *registry, uint64_t max_region)
    : stream_(stream), registry_(registry), budget_(budget), budget_zero_(budget >> 6), max_region_(max_region)
{
}

MemoryManager::~MemoryManager()
{
    for (auto &r : regions_) {
        if (registry_) registry_->remove(r.base);
        (void)hipFree(r.base);
    }
    for (auto &r : zero_regions_) {
        if (registry_) registry_->remove(r.base);
        (void)hipFree(r.base);
    }
}

void *MemoryManager::raw(uint64_t bytes, bool zero)
{
    bytes = round_up(bytes == 0 ? 1 : bytes);
    auto &regions = zero ? zero_regions_ : regions_;
    for (auto &r : regions) {
        if (r.size - r.used >= bytes) {
            void *p = r.base + r.used;
            r.used += bytes;
            return p;
        }
    }
    uint64_t &reserved = zero ? reserved_zero_ : reserved_;
    const uint64_t budget = zero ? budget_zero_ : budget_;
    const uint64_t remaining = budget > reserved ? budget - reserved : 0;
    if (remaining < bytes) throw std::bad_alloc();
    const uint64_t size = std::min(remaining, std::max(bytes, max_region_));
    Region r{nullptr, size, 0};
    SDB_HIP(hipMalloc(reinterpret_cast<void **>(&r.base), size));
    if (registry_) registry_->add(r.base, size);
    if (zero) {
        // zeroed once here, in stream order before any use; reset()
        // re-zeroes the part handed out since
        launch_fill_u8(r.base, 0, (int64_t)size, stream_);
    }
    reserved += size;
    r.used = bytes;
    regions.push_back(r);
    return r.base;
}

void MemoryManager::reset(hipStream_t s)
{
    for (auto &r : regions_) r.used = 0;
    for (auto &r : zero_regions_) {
        if (r.used) launch_fill_u8(r.base, 0, (int64_t)r.used, s);
        r.used = 0;
    }
}

void *Engine::phantom(uint64_t bytes)
{
    // addresses above the 57-bit virtual range: never device memory
    const uint64_t at = 0xF000000000000000ull + capture.phantom_next;
    capture.phantom_next += (bytes + 255) & ~255ull;
    return reinterpret_cast<void *>(at);
}

uint64_t MemoryManager::used_bytes() const
{
    uint64_t u = 0;
    for (auto &r : regions_) u += r.used;
    for (auto &r : zero_regions_) u += r.used;
    return u;
}

} // namespace sdb
