// tse_alloc.h -- all-or-nothing allocation of a group of device buffers.
//
// Header-only and free of HIP types, so that the state logic can be built
// and checked with a host compiler and a stand-in allocator.
#pragma once

#include <cstddef>

namespace tse {

// Allocates two buffers as one step.  On success *a and *b hold them and
// true is returned; on failure both are null, anything allocated on the way
// has been released, and false is returned.  alloc(void **p, size_t bytes)
// returns true on success; release(void *p) frees.
template <class Alloc, class Release>
bool alloc_both(void **a, size_t abytes, void **b, size_t bbytes, Alloc alloc,
                Release release) {
    *a = nullptr;
    *b = nullptr;
    void *pa = nullptr, *pb = nullptr;
    if (!alloc(&pa, abytes)) return false;
    if (!alloc(&pb, bbytes)) {
        release(pa);
        return false;
    }
    *a = pa;
    *b = pb;
    return true;
}

}  // namespace tse
