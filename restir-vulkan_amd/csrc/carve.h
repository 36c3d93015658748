// carve.h — one scratch allocation handed out as 256-byte-aligned pieces. A carve run with a null
// base only counts: end() is then the allocation size the same sequence of takes needs, so the
// size a context allocates and the offsets the kernels use come from one piece of code.
#pragma once

#include <cstddef>

namespace rk {

struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *b) : base(static_cast<char *>(b)) {}
    template <class T>
    T *take(size_t bytes) {
        T *r = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off = (off + bytes + 255) & ~static_cast<size_t>(255);
        return r;
    }
    size_t end() const { return off; }
};

}  // namespace rk
