// host_alloc.cpp - the library's own heap allocations, counted.
//
// Every operator new / delete the library's code calls binds to the
// definitions below: the link keeps them local to the shared object (only
// the scrs_* entry points are exported), so the process's other code keeps
// its own allocator and only this library's allocations are counted.  Both
// forward to malloc / free, as the default ones do, so memory may cross the
// boundary (a buffer from the C++ runtime freed here, or the reverse).
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <new>

namespace {

std::atomic<uint64_t> g_allocs{0};

void *counted_alloc(std::size_t n) {
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    return std::malloc(n ? n : 1);
}

void *counted_aligned(std::size_t n, std::size_t align) {
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    if (align < sizeof(void *)) align = sizeof(void *);
    void *p = nullptr;
    return posix_memalign(&p, align, n ? n : 1) == 0 ? p : nullptr;
}

}  // namespace

// Heap allocations made by the library's code since it was loaded.
extern "C" uint64_t scrs_host_alloc_count(void) { return g_allocs.load(std::memory_order_relaxed); }

void *operator new(std::size_t n) {
    if (void *p = counted_alloc(n)) return p;
    throw std::bad_alloc();
}
void *operator new[](std::size_t n) {
    if (void *p = counted_alloc(n)) return p;
    throw std::bad_alloc();
}
void *operator new(std::size_t n, const std::nothrow_t &) noexcept { return counted_alloc(n); }
void *operator new[](std::size_t n, const std::nothrow_t &) noexcept { return counted_alloc(n); }
void *operator new(std::size_t n, std::align_val_t a) {
    if (void *p = counted_aligned(n, static_cast<std::size_t>(a))) return p;
    throw std::bad_alloc();
}
void *operator new[](std::size_t n, std::align_val_t a) {
    if (void *p = counted_aligned(n, static_cast<std::size_t>(a))) return p;
    throw std::bad_alloc();
}
void *operator new(std::size_t n, std::align_val_t a, const std::nothrow_t &) noexcept {
    return counted_aligned(n, static_cast<std::size_t>(a));
}
void *operator new[](std::size_t n, std::align_val_t a, const std::nothrow_t &) noexcept {
    return counted_aligned(n, static_cast<std::size_t>(a));
}

void operator delete(void *p) noexcept { std::free(p); }
void operator delete[](void *p) noexcept { std::free(p); }
void operator delete(void *p, std::size_t) noexcept { std::free(p); }
void operator delete[](void *p, std::size_t) noexcept { std::free(p); }
void operator delete(void *p, const std::nothrow_t &) noexcept { std::free(p); }
void operator delete[](void *p, const std::nothrow_t &) noexcept { std::free(p); }
void operator delete(void *p, std::align_val_t) noexcept { std::free(p); }
void operator delete[](void *p, std::align_val_t) noexcept { std::free(p); }
void operator delete(void *p, std::size_t, std::align_val_t) noexcept { std::free(p); }
void operator delete[](void *p, std::size_t, std::align_val_t) noexcept { std::free(p); }
