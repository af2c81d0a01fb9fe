// Shared plumbing for the gfx950 stereo library: error codes, the thread-local
// error message behind sv_last_error(), launch checking and small helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#define SV_OK 0
#define SV_EINVAL -1        // bad argument (sizes, null pointers, ranges)
#define SV_EUNSUPPORTED -2  // valid for the reference, outside what the GPU path supports
#define SV_EHIP -3          // a HIP runtime error
#define SV_EWORKSPACE -4    // workspace too small

// element types of image inputs
#define SV_U8 0
#define SV_F32 1
#define SV_F64 2

// Tuning knobs: kernel variants that compute identical results at
// different speeds, chosen at run time (0 = the library's own heuristic).
#define SV_TUNE_SGM_PREFETCH 0  // steps of cost prefetch in the vertical/diagonal SGM kernel: 2, 4, 8
#define SV_TUNE_SGM_HTILE 1     // columns per LDS tile in the horizontal SGM kernel: 8, 16
#define SV_TUNE_CONV_ROWS 2     // output rows per thread in the tiled convolution: 1, 4
#define SV_TUNE_CONV_TW 3       // output columns per tile of the tiled convolution: 16, 32, 64
#define SV_TUNE_CONV_WSRC 4     // weights of the tiled convolution: 1 LDS, 2 scalar cache
#define SV_TUNE_SGM_OVERLAP 5   // horizontal SGM paths: 1 after the others, 2 beside them on a second stream
#define SV_TUNE_CONV_ENGINE 6   // tiled convolution: 1 vector ALUs, 2 matrix cores (FP32 4x4x1 MFMA; 3D layers slide over depth), 3 matrix cores, one depth per tile, 4 vector ALUs, one channel per staging pass, 5 matrix cores over 4-channel blocks with the next tile prefetched (2D), 6 matrix cores over a column of output depths per tile (3D), 7 direct, one output position per thread, no LDS (1 to 4 output channels)
#define SV_TUNE_CONV_MGROUPS 7  // matrix-core convolution: 64-pixel groups per wave, 1, 2, 4
#define SV_TUNE_SGM_SPLIT 8     // packed SGM kernels: 1 one lane per scanline, 2 two lanes per scanline
#define SV_TUNE_WTA_READS 9     // band epilogue (K = 21, u8): 1 one lane per pixel, 2 coalesced runs through LDS
#define SV_TUNE_CONV_DCHUNK 10  // depth-column 3D convolution: output depths per tile, 4, 6, 12
#define SV_TUNE_SGM_ORDER 11   // vertical/diagonal SGM blocks: 1 in launch order across the XCDs, 2 one contiguous run of (pair, direction, columns) per XCD
#define SV_TUNE_SGM_QUEUE 12   // packed SGM kernels with two lanes per scanline: 1 one workgroup per block of scanlines, 2 persistent workgroups pulling (image, direction, scanline block) items from per-XCD queues, longest scanlines first, 3 the same with one queue for the whole device
#define SV_TUNE_COUNT 13

namespace sv {

int set_error(int code, const char* fmt, ...);

// current value of a tuning knob (0: heuristic)
int tune(int knob);

// compute units of the current device (cached per device)
int cu_count();

inline size_t align_up(size_t n, size_t a = 256) { return (n + a - 1) / a * a; }

// Carves aligned sub-buffers out of one caller-provided workspace.
struct Carver {
    char* base;
    size_t used = 0;
    explicit Carver(void* p) : base(static_cast<char*>(p)) {}
    template <typename T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
        used += align_up(count * sizeof(T));
        return p;
    }
};

}  // namespace sv

#define SV_CHECK_LAUNCH()                                                          \
    do {                                                                           \
        hipError_t e_ = hipGetLastError();                                         \
        if (e_ != hipSuccess)                                                      \
            return sv::set_error(SV_EHIP, "%s: %s", __func__, hipGetErrorString(e_)); \
    } while (0)

#define SV_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) return sv::set_error(SV_EINVAL, __VA_ARGS__); \
    } while (0)
