// TW.h -- basic types of the paDIC engine (replaces TW/TW_Engine/TW.h:39-70).
// real_t is float as in the reference build; the GPU path computes in FP32
// with FP64 sums and solves where stated in the kernels.
#ifndef TWAMD_TW_H
#define TWAMD_TW_H

#include <hip/hip_vector_types.h>   // uchar1, float2, float4 for the public handle members

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>

#define TW_PI 3.14159265358979323846
#define TW_TWOPI 6.28318530717958647692
#define BLOCK_SIZE_256 256
#define BLOCK_SIZE_128 128
#define BLOCK_SIZE_64 64
#define TW_LIB_DLL_EXPORTS __attribute__((visibility("default")))

namespace TW {

using intensity_t = float;
using real_t = float;
using real_t2 = float2;
using real_t4 = float4;
using int_t = int;
using uint_t = unsigned int;

enum PARALLEL_COMPUTING_TYPE { Singlecore, Multicore, CUDA_GPU };

// cv::Mat-shaped view of an 8-bit single-channel image: rows, cols, step
// (bytes per row) and data.  Converts implicitly from any type with those
// members (a cv::Mat included), so callers written against the reference's
// `const cv::Mat&` parameters compile unchanged.
struct MatView {
    int rows = 0, cols = 0;
    size_t step = 0;
    const unsigned char* data = nullptr;
    MatView() = default;
    MatView(const unsigned char* d, int r, int c, size_t s) : rows(r), cols(c), step(s), data(d) {}
    template <class M, class = decltype(std::declval<const M&>().rows), class = decltype(std::declval<const M&>().data)>
    MatView(const M& m) : rows(m.rows), cols(m.cols), step((size_t)m.step), data((const unsigned char*)m.data) {}
};

// hcreateptr / hdestroyptr (TW_MemManager.cpp:3-96): one zeroed contiguous
// block plus row-pointer arrays, so p[0] (2-D) or p[0][0] (3-D) addresses the
// whole block.
template <typename T>
void hcreateptr(T*& p, size_t n)
{
    p = static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
    if (!p) throw std::bad_alloc();
}
template <typename T>
void hcreateptr(T**& p, size_t rows, size_t cols)
{
    T* block = static_cast<T*>(std::calloc(rows * cols ? rows * cols : 1, sizeof(T)));
    p = static_cast<T**>(std::malloc((rows ? rows : 1) * sizeof(T*)));
    if (!block || !p) throw std::bad_alloc();
    for (size_t i = 0; i < rows; ++i) p[i] = block + i * cols;
}
template <typename T>
void hcreateptr(T***& p, size_t d0, size_t d1, size_t d2)
{
    T* block = static_cast<T*>(std::calloc(d0 * d1 * d2 ? d0 * d1 * d2 : 1, sizeof(T)));
    T** rows = static_cast<T**>(std::malloc((d0 * d1 ? d0 * d1 : 1) * sizeof(T*)));
    p = static_cast<T***>(std::malloc((d0 ? d0 : 1) * sizeof(T**)));
    if (!block || !rows || !p) throw std::bad_alloc();
    for (size_t i = 0; i < d0 * d1; ++i) rows[i] = block + i * d2;
    for (size_t i = 0; i < d0; ++i) p[i] = rows + i * d1;
}
template <typename T>
void hdestroyptr(T*& p)
{
    std::free(p);
    p = nullptr;
}
template <typename T>
void hdestroyptr(T**& p)
{
    if (p) std::free(p[0]);
    std::free(p);
    p = nullptr;
}
template <typename T>
void hdestroyptr(T***& p)
{
    if (p) {
        std::free(p[0][0]);
        std::free(p[0]);
    }
    std::free(p);
    p = nullptr;
}

}  // namespace TW

#endif
