/*
 * memory.hip - the per-context caching device allocator (see Pool in
 * bcmg_internal.h).
 */
#include "bcmg_internal.h"

namespace bcmg {

namespace {
thread_local Pool *g_pool = nullptr;
}

Pool *current_pool() { return g_pool; }

PoolScope::PoolScope(Pool *p) : prev(g_pool) { g_pool = p; }
PoolScope::~PoolScope() { g_pool = prev; }

size_t Pool::size_class(size_t bytes)
{
    if (bytes <= 4096) return (bytes + 255) & ~(size_t)255;
    int e = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (e - 2);
    return (bytes + step - 1) & ~(step - 1);
}

void *Pool::alloc(size_t bytes)
{
    const size_t cls = size_class(bytes);
    void *p = nullptr;
    auto it = free_.find(cls);
    if (it != free_.end()) {
        p = it->second;
        free_.erase(it);
        cached -= cls;
    } else {
        if (hipMalloc(&p, cls) != hipSuccess) {
            (void)hipGetLastError();
            trim();
            if (hipMalloc(&p, cls) != hipSuccess) {
                (void)hipGetLastError();
                throw Error(BCMG_ENOMEM, "device allocation of " + std::to_string(cls) +
                                             " bytes failed");
            }
        }
    }
    cls_[p] = cls;
    live += cls;
    return p;
}

void Pool::release(void *p)
{
    auto it = cls_.find(p);
    if (it == cls_.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t cls = it->second;
    cls_.erase(it);
    live -= cls;
    if (free_released) {
        (void)hipFree(p);
        return;
    }
    free_.emplace(cls, p);
    cached += cls;
}

void Pool::trim()
{
    if (!free_.empty()) (void)hipDeviceSynchronize();
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached = 0;
}

Pool::~Pool() { trim(); }

void *dev_alloc(Pool *pool, size_t bytes)
{
    if (pool) return pool->alloc(bytes);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(BCMG_ENOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return p;
}

void dev_free(Pool *pool, void *p)
{
    if (!p) return;
    if (pool)
        pool->release(p);
    else
        (void)hipFree(p);
}

} // namespace bcmg
