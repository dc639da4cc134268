// k_book_o3.hip -- the contract-book kernel compiled for three waves per SIMD
// (amdgpu_waves_per_eu), one of the variants the book's autotuner times.
#include "k_book.inl"

namespace fdmc {

template <int MK, int GM>
struct BookO3 {
    static const void* fn() { return (const void*)&k_book_o3<MK, GM>; }
};

const void* book_kernel_o3(int mk, int gm) { return book_kernel_any<BookO3>(mk, gm); }

}  // namespace fdmc
