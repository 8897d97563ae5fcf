// stream_probe.hip -- the HBM bandwidth this GPU sustains for plain streams,
// the practical ceiling beside the 8 TB/s peak that the solver's kernels
// are priced against: a 16-byte-per-lane copy (half reads, half writes), a
// read-only pass and a write-only pass over two GiB-sized buffers, each
// with non-temporal loads and stores, grid-stride, timed with HIP events
// over 10 launches per grid size.  Prints one line per (pass, grid) and a
// JSON line with the best of each pass.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));

#define PROBE_CHECK(x)                                                                \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            std::fprintf(stderr, "stream_probe: %s: %s\n", #x, hipGetErrorString(e_)); \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

__global__ void __launch_bounds__(256) stream_copy(const f4 *__restrict__ x, f4 *__restrict__ y, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(x + i), y + i);
}

__global__ void __launch_bounds__(256) stream_read(const f4 *__restrict__ x, float *sink, size_t n) {
    float s = 0.f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const f4 v = __builtin_nontemporal_load(x + i);
        s += v.x + v.y + v.z + v.w;
    }
    if (s == -1.f) sink[0] = s;  // never true for the zeros read: keeps the loads
}

__global__ void __launch_bounds__(256) stream_write(f4 *__restrict__ y, size_t n) {
    const f4 one = {1.f, 1.f, 1.f, 1.f};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(one, y + i);
}

int main(int argc, char **argv) {
    const size_t gib = argc > 1 ? (size_t)std::atoi(argv[1]) : 4;
    if (gib < 1 || gib > 64) {
        std::fprintf(stderr, "stream_probe: buffer size must be 1..64 GiB\n");
        return 2;
    }
    const size_t bytes = gib << 30, n = bytes / sizeof(f4);
    f4 *x = nullptr, *y = nullptr;
    float *sink = nullptr;
    PROBE_CHECK(hipMalloc(&x, bytes));
    PROBE_CHECK(hipMalloc(&y, bytes));
    PROBE_CHECK(hipMalloc(&sink, 64));
    PROBE_CHECK(hipMemset(x, 0, bytes));
    PROBE_CHECK(hipMemset(y, 0, bytes));
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    const char *names[3] = {"copy", "read", "write"};
    double best[3] = {0.0, 0.0, 0.0};
    const int grids[] = {2048, 4096, 8192, 16384, 65536};
    for (int grid : grids) {
        for (int k = 0; k < 3; k++) {
            auto launch = [&]() {
                if (k == 0) hipLaunchKernelGGL(stream_copy, dim3(grid), dim3(256), 0, 0, x, y, n);
                if (k == 1) hipLaunchKernelGGL(stream_read, dim3(grid), dim3(256), 0, 0, x, sink, n);
                if (k == 2) hipLaunchKernelGGL(stream_write, dim3(grid), dim3(256), 0, 0, y, n);
            };
            launch();
            launch();
            const int reps = 10;
            PROBE_CHECK(hipEventRecord(e0));
            for (int r = 0; r < reps; r++) launch();
            PROBE_CHECK(hipEventRecord(e1));
            PROBE_CHECK(hipEventSynchronize(e1));
            PROBE_CHECK(hipGetLastError());
            float ms = 0.f;
            PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            const double moved = k == 0 ? 2.0 * (double)bytes : (double)bytes;
            const double gbs = moved / ms / 1e6;
            if (gbs > best[k]) best[k] = gbs;
            std::printf("[stream] %-5s grid %6d x 256: %.3f ms per pass, %.0f GB/s\n", names[k], grid, ms, gbs);
            std::fflush(stdout);
        }
    }
    std::printf("{\"buffer_gib\": %zu, \"copy_gbs\": %.0f, \"read_gbs\": %.0f, \"write_gbs\": %.0f}\n", gib, best[0],
                best[1], best[2]);
    (void)hipFree(x);
    (void)hipFree(y);
    (void)hipFree(sink);
    return 0;
}
